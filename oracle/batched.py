"""CPU restatement of the batched factor-and-solve (TEST INFRASTRUCTURE).

Per matrix the semantics are LAPACK dpotrf (lower) + dpotrs, as the
reference's one-tile dplasma_dposv runs them (src/zposv_wrapper.c:63-113,
src/cores/core_zpotrf.c:68-74 calling LAPACKE_dpotrf_work): on success L
overwrites the lower triangle and x = (L L^T)^-1 b; on failure info is the
1-based column of the first pivot that is not positive, the matrix is left
partly factored and no solve takes place.  Only tests/ import this module.

Also restated: the batch size sequence n_i = lo + splitmix64_i(seed) mod
(hi - lo + 1) (Steele, Lea & Flood 2014; word i of the generator started
at seed) and the batch layout (matrix i column-major at offset sum n_j^2,
right-hand side at offset sum n_j).

Parity status: pinned (tests/test_oracle_batched.py): numpy's Cholesky and
solve, exact factors of L0 L0^T with small-integer L0, info at the first,
a middle and the last column, and the published first splitmix64 outputs.
"""
import numpy as np

MASK = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15


def splitmix64(x):
    z = (x + GOLDEN) & MASK
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


def batch_sizes(seed, count, lo, hi):
    return [lo + splitmix64((seed + i * GOLDEN) & MASK) % (hi - lo + 1) for i in range(count)]


def offsets(ns):
    ao, bo, a, b = [], [], 0, 0
    for n in ns:
        ao.append(a)
        bo.append(b)
        a += n * n
        b += n
    return ao, bo, a, b


def potrf(A):
    """Unblocked right-looking Cholesky, lower, LAPACK info convention.
    Returns (L, info); L is the partly factored lower triangle on failure."""
    A = np.tril(np.array(A, dtype=np.float64))
    n = A.shape[0]
    for k in range(n):
        d = A[k, k]
        if not d > 0.0:
            return A, k + 1
        d = np.sqrt(d)
        A[k, k] = d
        A[k + 1:, k] /= d
        A[k + 1:, k + 1:] -= np.tril(np.outer(A[k + 1:, k], A[k + 1:, k]))
    return A, 0


def posv(A, b):
    """(L, x, info) of one system."""
    L, info = potrf(A)
    if info:
        return L, None, info
    y = np.linalg.solve(L, b)           # L is triangular: exact solve semantics
    x = np.linalg.solve(L.T, y)
    return L, x, 0
