"""Problem generators: the S_z-sector spin chain (drivers/matfuncs/SpinChainSZ.c)
and the honeycomb sheet (drivers/matfuncs/graphene.c).

Pins (CPU):
* the oracle's restatement of SpinChainSZ is checked entry by entry against a
  Hamiltonian built independently from Pauli matrices (Kronecker products,
  restricted to the sector, basis words ascending) and against the known
  ground-state energy of the 4-site ring; the library's generator (a
  different indexing scheme) is then compared with the oracle row by row,
  bit for bit;
* the oracle's honeycomb tree and the library's generator agree row by row;
  the nearest-neighbour sheet on a torus reproduces the analytic spectrum
  +-|1 + exp(2i k_w) + exp(i (k_l + k_w))| of the brick-wall lattice, the open
  sheet is bipartite (spectrum symmetric about 0), and the long-range sheet
  is symmetric.  The long-range values (t2, t3) are pinned only by the
  restatement cross-check and symmetry (parity partial).
* the on-site disorder differs from the reference by design (the reference
  draws C rand() in call order, which depends on how rows are visited); the
  library's values are checked for range, determinism and seed dependence.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol
import phist_amd as pa


def sorted_row(c, v):
    o = np.lexsort((v, c))
    return c[o], v[o]


def pauli_chain(L, obc=0):
    sx = np.array([[0, 1], [1, 0]], dtype=complex)
    sy = np.array([[0, -1j], [1j, 0]])
    sz = np.array([[1, 0], [0, -1]], dtype=complex)
    eye = np.eye(2)

    def site_op(op, i):
        # basis index sum_b bit_b 2^b: site 0 is the least significant factor
        mats = [op if j == i else eye for j in range(L)][::-1]
        out = mats[0]
        for m in mats[1:]:
            out = np.kron(out, m)
        return out

    H = np.zeros((2 ** L, 2 ** L), dtype=complex)
    for b in range(L - obc):
        c = (b + 1) % L
        for op in (sx, sy, sz):
            H += site_op(op, b) @ site_op(op, c)
    assert np.allclose(H.imag, 0)
    return H.real


@pytest.mark.parametrize("L,nup,obc", [(4, 2, 0), (6, 3, 0), (6, 2, 0), (8, 4, 0), (8, 4, 1), (10, 5, 0), (10, 3, 1)])
def test_spin_oracle_vs_pauli(L, nup, obc):
    rows = ol.spin_sz_rows(L, nup, obc)
    words = [s for s in range(2 ** L) if bin(s).count("1") == nup]
    assert len(rows) == len(words)
    H = pauli_chain(L, obc)[np.ix_(words, words)]
    A = ol.rows_to_dense(rows, len(words))
    assert np.array_equal(A, H)
    # a zero diagonal is omitted, never stored
    for r, (c, v) in enumerate(rows):
        assert np.all(v != 0.0)
        assert len(set(c.tolist())) == len(c)


def test_spin_ring4_ground_state():
    # sum of sigma.sigma over the 4-site ring: E0 = 4 * (-2) = -8 (total spin 0 singlet)
    A = ol.rows_to_dense(ol.spin_sz_rows(4, 2, 0), 6)
    assert np.isclose(np.linalg.eigvalsh(A)[0], -8.0)


@pytest.mark.parametrize("L,nup,obc", [(4, 2, 0), (8, 4, 0), (10, 4, 1), (12, 6, 0), (14, 7, 0), (16, 8, 1)])
def test_spin_library_matches_oracle(L, nup, obc):
    prm = pa.SpinSzParams(L, nup, obc, 0, 1.0, 1.0)
    n, mx = pa.generator_dims("spin_sz", prm)
    rows = ol.spin_sz_rows(L, nup, obc)
    assert n == len(rows) and mx == L + 1
    for r in range(n):
        gc, gv = sorted_row(*pa.generator_row("spin_sz", prm, r))
        wc, wv = sorted_row(*rows[r])
        assert np.array_equal(gc, wc) and np.array_equal(gv, wv), r
        fc, fv = sorted_row(*pa.generator_row("spin_sz", prm, r, np.float32))
        assert np.array_equal(fc, wc) and np.array_equal(fv, wv.astype(np.float32))


def test_spin_L24_sampled_rows():
    """the JD configuration: L = 24, S_z = 0, C(24,12) = 2,704,156 rows"""
    prm = pa.SpinSzParams(24, 12, 0, 0, 1.0, 1.0)
    n, mx = pa.generator_dims("spin_sz", prm)
    assert n == 2704156 and mx == 25
    # the oracle's tables for L = 24 are small (2^12 entries); rows sampled
    o = ol.oracle()
    import ctypes
    o.oracle_spin_sz_init.restype = ctypes.c_void_p
    nn = ctypes.c_int64(0)
    h = o.oracle_spin_sz_init(24, 12, 0, ctypes.byref(nn))
    assert nn.value == n
    cols, vals, k = np.zeros(25, np.int64), np.zeros(25), ctypes.c_int(0)
    rng = np.random.default_rng(24)
    for r in [0, 1, n - 1, n // 2] + list(rng.integers(0, n, 200)):
        o.oracle_spin_sz_row(ctypes.c_void_p(h), ctypes.c_int64(int(r)), ctypes.byref(k), ol.p(cols), ol.p(vals))
        gc, gv = sorted_row(*pa.generator_row("spin_sz", prm, int(r)))
        wc, wv = sorted_row(cols[:k.value].copy(), vals[:k.value].copy())
        assert np.array_equal(gc, wc) and np.array_equal(gv, wv)
    o.oracle_spin_sz_free(ctypes.c_void_p(h))


@pytest.mark.parametrize("W,L", [(4, 4), (6, 8), (5, 7), (8, 6), (3, 3)])
@pytest.mark.parametrize("pbc_w,pbc_l", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("lr", [0, 1])
def test_honeycomb_library_matches_oracle(W, L, pbc_w, pbc_l, lr):
    if (pbc_w and W < 5) or (pbc_l and L < 3):
        pytest.skip("wrap-around neighbours coincide on this small sheet")
    prm = pa.HoneycombParams(W, L, pbc_w, pbc_l, lr, 0, 0.0, 0)
    n, mx = pa.generator_dims("honeycomb", prm)
    assert n == W * L and mx == (13 if lr else 4)
    rows = ol.honeycomb_rows(W, L, pbc_w, pbc_l, lr)
    for r in range(n):
        gc, gv = sorted_row(*pa.generator_row("honeycomb", prm, r))
        wc, wv = sorted_row(*rows[r])
        assert len(gc) <= mx
        assert np.array_equal(gc, wc) and np.array_equal(gv, wv), r
    # a periodic wrap across an odd extent breaks the (l + w) parity pattern of
    # the vertical bonds (graphene.c's size set-up rounds W and L up to even)
    if (not pbc_w or W % 2 == 0) and (not pbc_l or L % 2 == 0):
        A = ol.rows_to_dense(rows, n)
        assert np.array_equal(A, A.T)


@pytest.mark.parametrize("W,L", [(4, 4), (6, 4), (8, 6), (10, 8)])
def test_honeycomb_torus_spectrum(W, L):
    """nearest-neighbour sheet, periodic in both directions: the two-site cells
    (l, w), (l, w+1) with l+w even form the Bravais lattice spanned by (0, 2)
    and (1, 1); A couples to B in the same cell and the cells shifted by
    -(0, 2) and -(1, 1), so E = +-|1 + exp(2i k_w) + exp(i (k_l + k_w))|"""
    A = ol.rows_to_dense(ol.honeycomb_rows(W, L, 1, 1, 0), W * L)
    got = np.sort(np.linalg.eigvalsh(A))
    ev = []
    for a, b in itertools.product(range(L), range(W // 2)):
        kl, kw = 2 * np.pi * a / L, 2 * np.pi * b / W
        e = abs(1 + np.exp(2j * kw) + np.exp(1j * (kl + kw)))
        ev += [e, -e]
    assert np.allclose(got, np.sort(ev), atol=1e-10)


@pytest.mark.parametrize("W,L", [(4, 4), (7, 5)])
def test_honeycomb_open_bipartite(W, L):
    rows = ol.honeycomb_rows(W, L, 0, 0, 0)
    A = ol.rows_to_dense(rows, W * L)
    ev = np.sort(np.linalg.eigvalsh(A))
    assert np.allclose(ev, -ev[::-1], atol=1e-10)
    deg = [int(np.sum(c != r)) for r, (c, _) in enumerate(rows)]
    assert max(deg) == 3 and min(deg) >= 1


def test_honeycomb_disorder():
    W, L = 20, 10
    p1 = pa.HoneycombParams(W, L, 0, 0, 0, 0, 1.0, 7)
    p2 = pa.HoneycombParams(W, L, 0, 0, 0, 0, 1.0, 8)
    d1 = np.array([dict(zip(*pa.generator_row("honeycomb", p1, r)))[r] for r in range(W * L)])
    d1b = np.array([dict(zip(*pa.generator_row("honeycomb", p1, r)))[r] for r in reversed(range(W * L))])[::-1]
    d2 = np.array([dict(zip(*pa.generator_row("honeycomb", p2, r)))[r] for r in range(W * L)])
    assert np.array_equal(d1, d1b)              # independent of the visiting order
    assert np.all((d1 >= -0.5) & (d1 < 0.5))    # gamma * (u - 1/2), u in [0, 1)
    assert not np.array_equal(d1, d2)
    assert abs(d1.mean()) < 0.1
