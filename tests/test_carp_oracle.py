"""CPU checks of the Kaczmarz (CARP) oracle, oracle/carp_oracle.c.

The reference's own CARP tests (test/kernels/CarpTest_def.hpp) pin little:
Identity_yields_zero (:600-617) and determinism (:470-560); its symmetry
tests are disabled.  The rest is pinned here by what the mathematics fixes:

* one row update with omega = 1 is the orthogonal projection onto the row's
  hyperplane, so that row's residual of (A - sigma I) x = b becomes zero;
* with b = 0 a forward + backward sweep is x -> M^H M x with M the product of
  the relaxed row projections, a Hermitian (real: symmetric) operator;
* the exact solution is a fixed point, and repeated sweeps with 0 < omega < 2
  converge to it (real and complex shifts);
* the RC variant agrees with the textbook complex Kaczmarz step
  x <- x - omega conj(r) (r^T x - b) / |r|^2, r = row of A - sigma I;
* rows of one colour of a distance-2 colouring commute: any order inside the
  colours gives the same bits, which is what lets the GPU run them at once.
"""
import numpy as np
import pytest

from oracle_lib import kacz_rows

RNG = np.random.default_rng(20260)


def rand_crs(n, per_row, diag=4.0, seed=0, sym_pattern=False):
    """random sparse matrix, 1-based builtin CRS with sorted columns and a diagonal"""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        cols = set(rng.choice(n, size=min(per_row, n), replace=False).tolist())
        cols.add(i)
        rows.append(sorted(cols))
    if sym_pattern:
        sets = [set(r) for r in rows]
        for i, r in enumerate(rows):
            for j in r:
                sets[j].add(i)
        rows = [sorted(s) for s in sets]
    ro = np.zeros(n + 1, np.int64)
    ci, va = [], []
    for i, r in enumerate(rows):
        for j in r:
            ci.append(j + 1)
            va.append(diag if j == i else rng.uniform(-1, 1))
        ro[i + 1] = len(ci)
    return ro + 1, np.array(ci, np.int32), np.array(va)


def dense(ro, ci, va, n):
    a = np.zeros((n, n))
    for i in range(n):
        for k in range(ro[i] - 1, ro[i + 1] - 1):
            a[i, ci[k] - 1] += va[k]
    return a


def sweep_seq(n):
    return np.concatenate([np.arange(n), np.arange(n)[::-1]]).astype(np.int32)


def greedy_colors(ro, ci, n):
    """distance-2 colouring: rows sharing a column (own index included) differ"""
    supp = [set((ci[ro[i] - 1:ro[i + 1] - 1] - 1).tolist()) | {i} for i in range(n)]
    touch = [[] for _ in range(n)]
    for i, s in enumerate(supp):
        for j in s:
            touch[j].append(i)
    color = [-1] * n
    for i in range(n):
        used = {color[r] for j in supp[i] for r in touch[j] if color[r] >= 0}
        c = 0
        while c in used:
            c += 1
        color[i] = c
    return np.array(color)


def test_identity_yields_zero():
    """CarpTest_def.hpp:600-617: A = I, sigma 0, omega 1, b 0 -> x = 0"""
    n = 40
    ro = np.arange(n + 1, dtype=np.int64) + 1
    ci = np.arange(1, n + 1, dtype=np.int32)
    va = np.ones(n)
    x = RNG.standard_normal((n, 3))
    kacz_rows(ro, ci, va, 0.0, 1.0, x, sweep_seq(n))
    assert np.all(x == 0.0)


@pytest.mark.parametrize("sigma", [0.0, 0.7, -1.3])
def test_row_projection_real(sigma):
    n = 30
    ro, ci, va = rand_crs(n, 5, seed=1)
    a = dense(ro, ci, va, n) - sigma * np.eye(n)
    x = RNG.standard_normal((n, 2))
    b = RNG.standard_normal((n, 2))
    for i in (0, 7, n - 1):
        kacz_rows(ro, ci, va, sigma, 1.0, x, [i], b=b)
        r = a[i] @ x - b[i]
        assert np.all(np.abs(r) <= 1e-13 * (np.abs(a[i]) @ np.abs(x) + np.abs(b[i])))


def test_row_projection_rc_matches_complex_kaczmarz():
    n = 25
    ro, ci, va = rand_crs(n, 4, seed=2)
    sr, si, om = np.array([0.4, -0.9]), np.array([1.1, 0.3]), np.array([1.0, 1.6])
    a = dense(ro, ci, va, n)
    xr, xi = RNG.standard_normal((n, 2)), RNG.standard_normal((n, 2))
    br = RNG.standard_normal((n, 2))
    z = xr + 1j * xi
    seq = [3, 11, 0, 24, 11]
    kacz_rows(ro, ci, va, sr, om, xr, seq, b=br, si=si, xi=xi)
    for c in range(2):
        m = a - (sr[c] + 1j * si[c]) * np.eye(n)
        zc = z[:, c].copy()
        for i in seq:
            r = m[i]
            zc = zc - om[c] * np.conj(r) * (r @ zc - br[i, c]) / np.vdot(r, r).real
        np.testing.assert_allclose(xr[:, c], zc.real, rtol=0, atol=1e-13)
        np.testing.assert_allclose(xi[:, c], zc.imag, rtol=0, atol=1e-13)
    # omega = 1 zeroes the complex row residual
    xr2, xi2 = RNG.standard_normal((n, 1)), RNG.standard_normal((n, 1))
    kacz_rows(ro, ci, va, sr[:1], 1.0, xr2, [5], b=br[:, :1].copy(), si=si[:1], xi=xi2)
    m = a - (sr[0] + 1j * si[0]) * np.eye(n)
    assert abs(m[5] @ (xr2[:, 0] + 1j * xi2[:, 0]) - br[5, 0]) < 1e-13


def test_sweep_operator_symmetric():
    n = 24
    ro, ci, va = rand_crs(n, 4, seed=3)
    x = np.eye(n)
    kacz_rows(ro, ci, va, 0.35, 1.7, x, sweep_seq(n))
    np.testing.assert_allclose(x, x.T, rtol=0, atol=1e-14)
    w = np.linalg.eigvalsh(x)
    assert w.min() > -1e-13 and w.max() < 1 + 1e-13  # M^T M with ||M|| <= 1


def test_sweep_operator_hermitian_rc():
    n = 20
    ro, ci, va = rand_crs(n, 4, seed=4)
    out = []
    for part in (1.0, 1j):
        xr = np.eye(n) * np.real(part)
        xi = np.eye(n) * np.imag(part)
        kacz_rows(ro, ci, va, 0.5, 1.4, xr, sweep_seq(n), si=-0.8, xi=xi)
        out.append(xr + 1j * xi)
    s, s_i = out
    np.testing.assert_allclose(s_i, 1j * s, rtol=0, atol=1e-14)  # complex-linear: S(i e_k) = i S e_k
    np.testing.assert_allclose(s, s.conj().T, rtol=0, atol=1e-14)


@pytest.mark.parametrize("rc", [False, True])
def test_fixed_point_and_convergence(rc):
    n = 60
    ro, ci, va = rand_crs(n, 4, diag=6.0, seed=5)
    a = dense(ro, ci, va, n)
    sr = np.array([0.5, -0.25, 2.0])
    si = np.array([0.3, -1.0, 0.7]) if rc else np.zeros(3)
    b = RNG.standard_normal((n, 3))
    xs = np.stack([np.linalg.solve(a - (sr[c] + 1j * si[c]) * np.eye(n), b[:, c]) for c in range(3)], axis=1)
    # fixed point
    xr, xi = xs.real.copy(), xs.imag.copy()
    kacz_rows(ro, ci, va, sr, 1.2, xr, sweep_seq(n), b=b, si=si if rc else None, xi=xi if rc else None)
    assert np.abs(xr + 1j * xi - xs).max() < 1e-12 if rc else np.abs(xr - xs.real).max() < 1e-12
    # convergence from zero
    xr, xi = np.zeros((n, 3)), np.zeros((n, 3))
    for _ in range(200):
        kacz_rows(ro, ci, va, sr, 1.2, xr, sweep_seq(n), b=b, si=si if rc else None, xi=xi if rc else None)
    got = xr + 1j * xi if rc else xr
    assert np.abs(got - xs).max() < 1e-9 * np.abs(xs).max()


def test_b_none_equals_zero():
    n = 30
    ro, ci, va = rand_crs(n, 5, seed=6)
    x0 = RNG.standard_normal((n, 4))
    x1, x2 = x0.copy(), x0.copy()
    kacz_rows(ro, ci, va, [0.1, 0.2, 0.3, 0.4], 1.5, x1, sweep_seq(n))
    kacz_rows(ro, ci, va, [0.1, 0.2, 0.3, 0.4], 1.5, x2, sweep_seq(n), b=np.zeros((n, 4)))
    assert np.array_equal(x1, x2)


def test_rows_of_a_colour_commute():
    n = 80
    ro, ci, va = rand_crs(n, 3, seed=7, sym_pattern=True)
    color = greedy_colors(ro, ci, n)
    nc = color.max() + 1
    assert nc < n
    x0 = RNG.standard_normal((n, 2))
    b = RNG.standard_normal((n, 2))
    outs = []
    for trial in range(3):
        rng = np.random.default_rng(trial)
        fwd = np.concatenate([rng.permutation(np.flatnonzero(color == c)) for c in range(nc)])
        bwd = np.concatenate([rng.permutation(np.flatnonzero(color == c)) for c in reversed(range(nc))])
        x = x0.copy()
        kacz_rows(ro, ci, va, [0.3, -0.6], [1.0, 1.5], x, np.concatenate([fwd, bwd]), b=b)
        outs.append(x)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    # a different row order (not by colours) changes the bits
    x = x0.copy()
    kacz_rows(ro, ci, va, [0.3, -0.6], [1.0, 1.5], x, sweep_seq(n), b=b)
    assert not np.array_equal(x, outs[0])


def test_float_matches_double():
    n = 50
    ro, ci, va = rand_crs(n, 5, seed=8)
    x = RNG.standard_normal((n, 2))
    b = RNG.standard_normal((n, 2))
    xd = x.copy()
    kacz_rows(ro, ci, va, [0.2, 0.9], 1.3, xd, sweep_seq(n), b=b)
    xf = x.astype(np.float32)
    kacz_rows(ro, ci, va.astype(np.float32), np.float32([0.2, 0.9]), np.float32(1.3), xf, sweep_seq(n),
              b=b.astype(np.float32))
    np.testing.assert_allclose(xf, xd, rtol=0, atol=2e-5 * np.abs(xd).max())
