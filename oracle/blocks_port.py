"""CPU Gram of a tall-skinny block for the CPU references of the Gram and
CholQR benches.

TEST INFRASTRUCTURE AND CPU REFERENCE TIMING ONLY (bench.py's cpu_baseline leg and
tests/); the product library never imports, links or runs it.

The reference's builtin Gram (gemm_kernels_sC.F90:17-265: an OpenMP loop
over rows with a reduction of the k x k partial sums) needs PHIST's
CMake-generated header and does not build here.  This restates its shape in
numpy: rows split into one block per thread (a thread pool, each worker's
BLAS on one thread), each block's Gram by BLAS, the partial Grams added in
block order.  A BLAS Gram of a block narrower than 32 columns runs far below
the memory's rate, so rows are folded: a block of n x k (k = 8, 16) read as
n/f x 32 (f = 32/k consecutive rows side by side) has the Gram whose f
diagonal k x k blocks sum to the wanted one (the same products, added in
another order).
"""
import numpy as np

from minres_port import RowPool


def _gram_block(V, r0, r1):
    k = V.shape[1]
    f = 32 // k if k < 32 and 32 % k == 0 else 1
    if f == 1 or (r1 - r0) % f or not V.flags.c_contiguous:
        B = V[r0:r1]
        return B.T @ B
    W = V[r0:r1].reshape(-1, f * k)
    G = W.T @ W
    out = G[:k, :k].copy()
    for a in range(1, f):
        out += G[a * k:(a + 1) * k, a * k:(a + 1) * k]
    return out


def gram(V, threads):
    """V'V of an (n, k) row-major block on `threads` threads"""
    from threadpoolctl import threadpool_limits
    n = V.shape[0]
    blocks = max(1, threads)
    step = -(-n // blocks)
    step += (-step) % 32  # blocks of a whole number of folded rows
    pool = RowPool(n, blocks)
    pool.blocks = [(r, min(n, r + step)) for r in range(0, n, step)]
    try:
        with threadpool_limits(1, user_api="blas"):
            return pool.colsum(lambda r0, r1: _gram_block(V, r0, r1))
    finally:
        pool.close()
