"""Pins the oracle's Cholesky adjoint (oracle/adjoint.py) on the CPU:
central finite differences of numpy's Cholesky (brute force, small N), the
closed forms d logdet(A)/dA = A^-1 and d(b^T A^-1 b)/dA = -A^-1 b b^T A^-1,
the Gaussian-process gradient (K^-1 - alpha alpha^T)/2, and the blocked
(tile) form against the closed form on ragged tilings.  A planted mistake
(a transposed operand, a dropped diagonal halving) must fail."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import adjoint as oad  # noqa: E402


def spd(N, seed):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((N, N))
    return B @ B.T / N + 0.5 * np.eye(N)


@pytest.mark.parametrize("N", [1, 2, 5, 8])
def test_closed_form_vs_finite_differences(N):
    A = spd(N, N)
    W = np.tril(np.random.default_rng(N + 1).standard_normal((N, N)))
    f = lambda L: np.sum(W * L)                     # noqa: E731  df/dL = W
    got = oad.adjoint_closed(np.linalg.cholesky(A), W)
    fd = oad.fd_adjoint(A, f)
    assert np.abs(got - fd).max() < 1e-7 * max(1.0, np.abs(fd).max())
    assert np.array_equal(got, got.T)


def test_closed_forms():
    N = 40
    A = spd(N, 3)
    L = np.linalg.cholesky(A)
    Ai = np.linalg.inv(A)
    # logdet A = 2 sum log L_ii
    assert np.abs(oad.adjoint_closed(L, np.diag(2.0 / np.diag(L))) - Ai).max() < 1e-12 * np.abs(Ai).max()
    # b^T A^-1 b = |L^-1 b|^2: Lbar = -2 L^-T (L^-1 b)(L^-1 b)^T, lower part
    b = np.random.default_rng(0).standard_normal(N)
    z = np.linalg.solve(L, b)
    Lbar = np.tril(-2.0 * np.outer(np.linalg.solve(L.T, z), z))
    want = -np.outer(Ai @ b, Ai @ b)
    assert np.abs(oad.adjoint_closed(L, Lbar) - want).max() < 1e-11 * np.abs(want).max()


def test_gp_gradient():
    rng = np.random.default_rng(5)
    X = rng.uniform(size=(30, 3))
    K = oad.se_kernel(X, 0.2, 1.0, 1e-2)
    y = rng.standard_normal(30)
    G = oad.gp_nll_grad_K(K, y)
    # central differences over a symmetric pair and a diagonal entry
    for (i, j) in ((3, 7), (4, 4)):
        E = np.zeros_like(K)
        E[i, j] = E[j, i] = 1e-6
        d = (oad.gp_nll(K + E, y) - oad.gp_nll(K - E, y)) / 2e-6
        want = G[i, j] if i == j else 2 * G[i, j]
        assert abs(d - want) < 1e-5 * max(1.0, abs(want))


@pytest.mark.parametrize("N,NB", [(9, 2), (9, 4), (100, 17), (300, 64), (300, 93), (256, 256)])
def test_tile_form_matches_closed_form(N, NB):
    A = spd(N, NB)
    L = np.linalg.cholesky(A)
    W = np.tril(np.random.default_rng(2).standard_normal((N, N)))
    ref = oad.adjoint_closed(L, W)
    assert np.abs(oad.tile_adjoint(L, W, NB) - ref).max() < 1e-12 * N * np.abs(ref).max()


def test_planted_mistakes_fail():
    N = 6
    A = spd(N, 1)
    L = np.linalg.cholesky(A)
    W = np.tril(np.random.default_rng(3).standard_normal((N, N)))
    fd = oad.fd_adjoint(A, lambda L_: np.sum(W * L_))
    # Phi without halving the diagonal
    P = np.tril(L.T @ W)
    S = np.linalg.solve(L.T, np.linalg.solve(L.T, P).T).T
    assert np.abs(0.5 * (S + S.T) - fd).max() > 1e-4
    # transposed product inside Phi
    P = oad.phi(W.T @ L)
    S = np.linalg.solve(L.T, np.linalg.solve(L.T, P).T).T
    assert np.abs(0.5 * (S + S.T) - fd).max() > 1e-4
