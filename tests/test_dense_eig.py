"""CPU: the library's host symmetric eigensolvers (phist_amd_dense_sym_eig) against numpy.linalg.eigh,
and the invariants A V = V diag(w), V'V = I, on random, degenerate, diagonal, tridiagonal and
Rayleigh-Ritz-shaped (diagonal block bordered by dense columns) matrices.  No device work."""
import ctypes

import numpy as np
import pytest

import phist_amd


def eig(a, method):
    n = a.shape[0]
    f = np.asfortranarray(a.copy())
    w = np.zeros(max(1, n))
    fl = ctypes.c_int(0)
    phist_amd.lib().phist_amd_dense_sym_eig(f.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(n),
                                             ctypes.c_int(max(1, n)), w.ctypes.data_as(ctypes.c_void_p),
                                             ctypes.c_int(method), ctypes.byref(fl))
    assert fl.value == 0
    return w[:n], np.array(f)


def cases():
    rng = np.random.default_rng(7)
    out = []
    for n in (1, 2, 3, 5, 10, 33, 48, 64):
        b = rng.standard_normal((n, n))
        out.append((f"rand{n}", b + b.T))
    out.append(("identity8", np.eye(8)))
    out.append(("zero5", np.zeros((5, 5))))
    d = np.diag(rng.standard_normal(20))
    out.append(("diag20", d))
    t = np.diag(rng.standard_normal(30)) + np.diag(rng.standard_normal(29), 1)
    out.append(("tridiag30", t + np.triu(t, 1).T))
    rep = np.diag([1.0, 1.0, 1.0, 2.0, 2.0, 5.0])
    qq, _ = np.linalg.qr(rng.standard_normal((6, 6)))
    out.append(("repeated6", qq @ rep @ qq.T))
    h = np.diag(np.sort(rng.standard_normal(44)))  # locked/restarted part of a JD projection
    border = rng.standard_normal((48, 4)) * 1e-3
    hh = np.zeros((48, 48))
    hh[:44, :44] = h
    hh[:, 44:] = border
    hh[44:, :] = border.T
    hh[44:, 44:] = (border[44:] + border[44:].T)
    out.append(("ritz48", hh))
    out.append(("scaled", (lambda b: 1e8 * (b + b.T))(rng.standard_normal((16, 16)))))
    return out


@pytest.mark.parametrize("method", [0, 1])
@pytest.mark.parametrize("name,a", cases(), ids=[c[0] for c in cases()])
def test_sym_eig(method, name, a):
    n = a.shape[0]
    w, v = eig(a, method)
    ref = np.linalg.eigvalsh(a)
    scale = max(1.0, np.abs(ref).max()) if n else 1.0
    assert np.all(np.diff(w) >= 0)
    np.testing.assert_allclose(w, ref, rtol=0, atol=1e-13 * scale * max(1, n))
    if n:
        np.testing.assert_allclose(v.T @ v, np.eye(n), rtol=0, atol=1e-13 * n)
        np.testing.assert_allclose(a @ v, v * w, rtol=0, atol=1e-13 * scale * n)
