"""CPU restatement of the blocked Cholesky adjoint (TEST INFRASTRUCTURE).

The reference has no Cholesky adjoint: this is the differentiation growth
path of its factorization (src/zpotrf_L.jdf:91-633 run in reverse).  Only
tests/ import this module, as the checker of the GPU kernels.

Convention.  A is symmetric positive definite, L = chol(A) lower, f a scalar
function of L with gradient Lbar (lower triangular).  The adjoint returned is
the symmetric gradient Abar, i.e. for a symmetric perturbation dA

    df = sum_ij Abar_ij dA_ij        (Abar = Abar^T)

so perturbing the pair A_ij = A_ji by h changes f by 2 h Abar_ij (i != j).
This is the convention of torch.linalg.cholesky's backward.

Closed form (Murray 2016, "Differentiation of the Cholesky decomposition",
arXiv:1602.07527, eq. 10; Smith 1995):

    P    = Phi(L^T Lbar)       Phi: lower triangle, diagonal halved
    S    = L^-T P L^-1
    Abar = (S + S^T) / 2

Blocked form (tile_adjoint below; the GPU kernels follow its steps).  The
right-looking factorization is, per tile column k,
    L_kk = chol(A_kk),  L_mk = A_mk L_kk^-T,  A_mn -= L_mk L_nk^T  (m >= n > k)
Reversing it, with Abar of the trailing matrix already final, for
k = MT-1 .. 0:
    S1  Lbar_mk -= 2 sum_{n>k} Abar_(m,n) L_nk          (m > k; Abar full symmetric)
    S2  Abar_mk  = Lbar_mk L_kk^-1 / 2                  (m > k)
    S3  Lbar_kk -= tril(2 sum_{m>k} Abar_mk^T L_mk)
    S4  Abar_kk  = closed form on the tile (L_kk, Lbar_kk)
S1 is the transpose of the trailing update (its 2/3 N^3 flops dominate);
S2 and S3 reverse the panel TRSM; S4 reverses the diagonal POTRF.
Parity status: pinned (see tests/test_oracle_adjoint.py): finite differences
of numpy's Cholesky on small matrices, the closed forms
d logdet A / dA = A^-1 and d (b^T A^-1 b) / dA = -A^-1 b b^T A^-1, and the
blocked form against the closed form on ragged tilings.
"""
import numpy as np


def phi(X):
    """Lower triangle of X with the diagonal halved."""
    P = np.tril(X)
    P[np.diag_indices_from(P)] *= 0.5
    return P


def adjoint_closed(L, Lbar):
    """Symmetric gradient Abar of f(chol(A)) from L and Lbar (closed form)."""
    L = np.tril(L)
    Lbar = np.tril(Lbar)
    P = phi(L.T @ Lbar)
    # S = L^-T P L^-1 by two triangular solves
    X = np.linalg.solve(L.T, P)            # L^-T P
    S = np.linalg.solve(L.T, X.T).T        # (L^-T P) L^-1 = (L^-T (L^-T P)^T)^T
    return 0.5 * (S + S.T)


def tile_adjoint(L, Lbar, NB):
    """Blocked reverse of the right-looking tile Cholesky (steps S1-S4 of the
    module docstring) on dense N x N arrays; returns Abar (full symmetric)."""
    N = L.shape[0]
    L = np.tril(L).astype(np.float64)
    G = np.tril(Lbar).astype(np.float64)     # holds Lbar, overwritten by Abar
    Ab = np.zeros((N, N))                     # full symmetric Abar of finished columns
    MT = (N + NB - 1) // NB
    sl = [slice(t * NB, min(N, (t + 1) * NB)) for t in range(MT)]
    for k in range(MT - 1, -1, -1):
        kk, below = sl[k], slice(sl[k].stop, N)
        Lkk = L[kk, kk]
        if below.start < N:
            # S1: Lbar(:, k) below the diagonal -= 2 Abar_trailing L(:, k)
            G[below, kk] -= 2.0 * Ab[below, below] @ L[below, kk]
            # S2: Abar_mk = Lbar_mk L_kk^-1 / 2
            Ab[below, kk] = 0.5 * np.linalg.solve(Lkk.T, G[below, kk].T).T
            Ab[kk, below] = Ab[below, kk].T
            # S3: Lbar_kk -= tril(2 Abar_mk^T L_mk)
            G[kk, kk] -= np.tril(2.0 * Ab[below, kk].T @ L[below, kk])
        # S4: the diagonal tile
        Ab[kk, kk] = adjoint_closed(Lkk, G[kk, kk])
    return Ab


def fd_adjoint(A, f, h=1e-6):
    """Central finite differences of f(chol(A)) over symmetric perturbations
    (brute force, small N): returns Abar with the convention above."""
    N = A.shape[0]
    Ab = np.zeros((N, N))
    for i in range(N):
        for j in range(i + 1):
            E = np.zeros((N, N))
            E[i, j] = E[j, i] = h
            d = (f(np.linalg.cholesky(A + E)) - f(np.linalg.cholesky(A - E))) / (2 * h)
            Ab[i, j] = Ab[j, i] = d if i == j else 0.5 * d
    return Ab


def se_kernel(X, lengthscale, variance, noise):
    """Squared-exponential kernel matrix K_ij = v exp(-|x_i - x_j|^2 / (2 l^2))
    + noise * delta_ij of the Gaussian-process likelihood."""
    d2 = np.sum((X[:, None, :] - X[None, :, :]) ** 2, axis=-1)
    return variance * np.exp(-0.5 * d2 / lengthscale ** 2) + noise * np.eye(X.shape[0])


def gp_nll(K, y):
    """Negative log marginal likelihood 1/2 y^T K^-1 y + 1/2 logdet K + N/2 log 2 pi."""
    L = np.linalg.cholesky(K)
    a = np.linalg.solve(L, y)
    return 0.5 * a @ a + np.sum(np.log(np.diag(L))) + 0.5 * len(y) * np.log(2 * np.pi)


def gp_nll_grad_K(K, y):
    """Analytic symmetric gradient of gp_nll w.r.t. K: (K^-1 - alpha alpha^T) / 2,
    alpha = K^-1 y."""
    Ki = np.linalg.inv(K)
    al = Ki @ y
    return 0.5 * (Ki - np.outer(al, al))
