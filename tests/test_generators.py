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
  sheet is bipartite (spectrum symmetric about 0).  The long-range sheet
  (second neighbours t2, third neighbours t3 across each hexagon, 3*t3 on the
  diagonal) equals, entry by entry, a Hamiltonian built here from the
  nearest-neighbour graph alone (second neighbours = graph distance 2, third
  neighbours = opposite corners of the graph's 6-cycles), on open, cylinder
  and torus sheets; on the torus its spectrum is the union of the 2x2 Bloch
  Hamiltonians H(k) over the cell momenta.
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


T1, T2, T3 = -1.0, -0.1 / 2.78, -0.095 / 2.78  # graphene.c:58-60


def brickwall_nn(W, L, pbc_w, pbc_l):
    """nearest-neighbour adjacency of the brick-wall honeycomb, site (l, w) = row l*W + w: every site is
    bonded to (l, w+-1); a site with l+w odd also to (l+1, w), with l+w even to (l-1, w)"""
    n = W * L
    adj = [set() for _ in range(n)]

    def site(l, w):
        if pbc_w:
            w %= W
        if pbc_l:
            l %= L
        return l * W + w if 0 <= l < L and 0 <= w < W else None

    for l in range(L):
        for w in range(W):
            a = l * W + w
            for b in (site(l, w + 1), site(l + 1, w) if (l + w) % 2 else None):
                if b is not None and b != a:
                    adj[a].add(b)
                    adj[b].add(a)
    return adj


def honeycomb_from_graph(adj):
    """t1 on graph neighbours, t2 on graph distance 2, t3 between opposite corners of 6-cycles
    (paths a-b-c-d with a 3-step return d-e-f-a through distinct vertices), 3*t3 on the diagonal"""
    n = len(adj)
    H = np.zeros((n, n))
    for a in range(n):
        H[a, a] = 3 * T3
        for b in adj[a]:
            H[a, b] = T1
        d2 = {c for b in adj[a] for c in adj[b]} - adj[a] - {a}
        for c in d2:
            H[a, c] = T2
        opp = set()
        for b in adj[a]:
            for c in adj[b] - {a}:
                for d in adj[c] - {a, b}:
                    if d in adj[a] or d in d2:
                        continue
                    # a 6-cycle a-b-c-d-e-f-a with e, f new vertices
                    for e in adj[d] - {a, b, c}:
                        for f in adj[e] - {a, b, c, d}:
                            if a in adj[f]:
                                opp.add(d)
        for d in opp:
            H[a, d] = T3
    return H


def honeycomb_sheet(W, L, pbc_w, pbc_l):
    """the long-range sheet from the graph of the infinite lattice: open directions are padded by two
    sites on both sides (an even shift keeps the sublattice pattern), the shells are found on the padded
    graph and restricted to the sheet, so a bond exists when its two sites do (graphene.c's edges)"""
    pw, pl = (0 if pbc_w else 2), (0 if pbc_l else 2)
    Wp, Lp = W + 2 * pw, L + 2 * pl
    Hp = honeycomb_from_graph(brickwall_nn(Wp, Lp, pbc_w, pbc_l))
    idx = np.array([(l + pl) * Wp + (w + pw) for l in range(L) for w in range(W)])
    return Hp[np.ix_(idx, idx)]


@pytest.mark.parametrize("W,L,pbc_w,pbc_l", [(6, 4, 0, 0), (8, 6, 0, 0), (7, 5, 0, 0), (8, 6, 1, 0), (8, 5, 1, 0),
                                             (6, 6, 0, 1), (7, 4, 0, 1), (8, 6, 1, 1), (10, 8, 1, 1)])
def test_honeycomb_long_range_from_graph(W, L, pbc_w, pbc_l):
    """periodic directions need an even extent (the sublattice pattern) and W >= 8 around the w ring
    (a 6-site ring would be a 6-cycle of its own)"""
    A = ol.rows_to_dense(ol.honeycomb_rows(W, L, pbc_w, pbc_l, 1), W * L)
    assert np.array_equal(A, honeycomb_sheet(W, L, pbc_w, pbc_l))


def test_honeycomb_long_range_catches_planted_errors():
    """the graph construction separates the three shells: swapping t2 and t3, or dropping the
    3*t3 diagonal, changes the matrix"""
    W, L = 8, 6
    A = ol.rows_to_dense(ol.honeycomb_rows(W, L, 1, 1, 1), W * L)
    H = honeycomb_sheet(W, L, 1, 1)
    swapped = np.where(H == T2, T3, np.where(H == T3, T2, H))
    assert not np.array_equal(A, swapped)
    assert not np.array_equal(A, H - 3 * T3 * np.eye(W * L))
    deg = [(np.sum(H[r] == T1), np.sum(H[r] == T2), np.sum(H[r] == T3)) for r in range(W * L)]
    assert set(deg) == {(3, 6, 3)}  # honeycomb shells: 3 first, 6 second, 3 third neighbours


@pytest.mark.parametrize("W,L", [(8, 6), (10, 8), (12, 4)])
def test_honeycomb_long_range_torus_bloch(W, L):
    """torus spectrum of the long-range sheet = eigenvalues of the 2x2 Bloch Hamiltonians
    H(k)_{ab} = sum over bonds from the a-site of a cell to b-sites of t * exp(i k.d) (d the bond vector,
    minimum image), k = (2 pi p / L, 2 pi q / W), q < W/2 (cells spanned by (0, 2) and (1, 1))"""
    A = ol.rows_to_dense(ol.honeycomb_rows(W, L, 1, 1, 1), W * L)
    got = np.sort(np.linalg.eigvalsh(A))
    # bonds of the two sites of the cell at the origin: A = (0, 0) (l+w even), B = (0, 1)
    H = honeycomb_sheet(W, L, 1, 1)
    sites = [(0, 0), (0, 1)]
    ev = []
    for pp, qq in itertools.product(range(L), range(W // 2)):
        k = np.array([2 * np.pi * pp / L, 2 * np.pi * qq / W])
        Hk = np.zeros((2, 2), complex)
        for ia, (la, wa) in enumerate(sites):
            ra = la * W + wa
            for rb in np.nonzero(H[ra])[0]:
                lb, wb = divmod(int(rb), W)
                dl = (lb - la + L // 2) % L - L // 2
                dw = (wb - wa + W // 2) % W - W // 2
                ib = (lb + wb) % 2
                Hk[ia, ib] += H[ra, rb] * np.exp(1j * (k[0] * dl + k[1] * dw))
        ev += list(np.linalg.eigvalsh(Hk))
    assert np.allclose(got, np.sort(ev), atol=1e-10)
