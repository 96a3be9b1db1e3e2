"""The differentiable Cholesky (dplasma_amd.autograd over the C ABI) on the
Gaussian-process negative log marginal likelihood: N points uniform in
[0,1]^3 (seed 5), squared-exponential kernel (lengthscale 0.2, variance 1),
noise 1e-2, y ~ N(0, K).  Checked against
  - the FP64 analytic gradient dNLL/dK = (K^-1 - alpha alpha^T) / 2,
  - central finite differences (h = 1e-6) over symmetric pairs of K entries
    and over the lengthscale,
  - the oracle's closed-form adjoint for an arbitrary Lbar.
FP64 tolerances are stated at each assert; the FP32 run (mixed precision:
FP32 factor and adjoint of the FP64 kernel matrix rounded to FP32) is held
to 2e-2 of the largest gradient entry."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import adjoint as oad  # noqa: E402

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ag(dp):
    import dplasma_amd.autograd as ag
    return ag


def gp_problem(N, seed=5):
    rng = np.random.default_rng(seed)
    X = rng.uniform(size=(N, 3))
    K = oad.se_kernel(X, 0.2, 1.0, 1e-2)
    y = np.linalg.cholesky(K) @ rng.standard_normal(N)
    return X, K, y


def test_cholesky_forward_and_adjoint(ag):
    N, NB = 700, 128
    rng = np.random.default_rng(1)
    B = rng.standard_normal((N, N))
    A = B @ B.T / N + 0.5 * np.eye(N)
    W = np.tril(rng.standard_normal((N, N)))
    At = torch.tensor(A, device="cuda", requires_grad=True)
    L = ag.cholesky(At, NB)
    Lr = np.linalg.cholesky(A)
    assert np.abs(L.detach().cpu().numpy() - Lr).max() <= 1e-13 * np.abs(Lr).max()
    (L * torch.tensor(W, device="cuda")).sum().backward()
    want = oad.adjoint_closed(Lr, W)
    got = At.grad.cpu().numpy()
    assert np.abs(got - want).max() <= 1e-12 * N * np.abs(want).max()
    assert np.array_equal(got, got.T)


@pytest.mark.parametrize("recompute", [False, True])
def test_gp_gradient_analytic_and_fd(ag, recompute):
    N, NB = 1024, 256
    X, K, y = gp_problem(N)
    Kt = torch.tensor(K, device="cuda", requires_grad=True)
    yt = torch.tensor(y, device="cuda")
    nll = ag.gp_nll(Kt, yt, NB, recompute)
    nll.backward()
    G = Kt.grad.cpu().numpy()
    # value against numpy
    assert abs(nll.item() - oad.gp_nll(K, y)) <= 1e-10 * abs(oad.gp_nll(K, y))
    # analytic gradient (K^-1 - alpha alpha^T) / 2
    Ga = oad.gp_nll_grad_K(K, y)
    assert np.abs(G - Ga).max() <= 1e-9 * np.abs(Ga).max(), np.abs(G - Ga).max()
    # central differences over symmetric pairs and diagonal entries, h = 1e-6
    h = 1e-6
    for (i, j) in ((0, 0), (5, 17), (100, 3), (511, 510), (1023, 0)):
        E = torch.zeros_like(Kt.detach())
        E[i, j] = E[j, i] = h
        with torch.no_grad():
            fp = ag.gp_nll(Kt.detach() + E, yt, NB).item()
            fm = ag.gp_nll(Kt.detach() - E, yt, NB).item()
        fd = (fp - fm) / (2 * h)
        want = G[i, j] if i == j else 2 * G[i, j]
        assert abs(fd - want) <= 1e-5 * max(1.0, abs(want)), ((i, j), fd, want)


def test_gp_lengthscale_gradient(ag):
    N, NB = 1024, 256
    X, K, y = gp_problem(N)
    Xt = torch.tensor(X, device="cuda")
    yt = torch.tensor(y, device="cuda")
    ell = torch.tensor(0.2, dtype=torch.float64, device="cuda", requires_grad=True)
    nll = ag.gp_nll(ag.se_kernel(Xt, ell, 1.0, 1e-2), yt, NB)
    nll.backward()
    h = 1e-6
    with torch.no_grad():
        fp = ag.gp_nll(ag.se_kernel(Xt, 0.2 + h, 1.0, 1e-2), yt, NB).item()
        fm = ag.gp_nll(ag.se_kernel(Xt, 0.2 - h, 1.0, 1e-2), yt, NB).item()
    fd = (fp - fm) / (2 * h)
    # analytic: sum_ij Kbar_ij dK_ij/dl, dK/dl = K_se * d2 / l^3
    d2 = np.sum((X[:, None, :] - X[None, :, :]) ** 2, axis=-1)
    dK = np.exp(-0.5 * d2 / 0.04) * d2 / 0.2 ** 3
    an = np.sum(oad.gp_nll_grad_K(K, y) * dK)
    assert abs(ell.grad.item() - an) <= 1e-8 * abs(an)
    assert abs(fd - an) <= 1e-5 * abs(an), (fd, an)


def test_gp_fp32_mixed(ag):
    N, NB = 1024, 256
    X, K, y = gp_problem(N)
    Kt = torch.tensor(K, device="cuda").float().requires_grad_(True)
    yt = torch.tensor(y, device="cuda").float()
    ag.gp_nll(Kt, yt, NB).backward()
    G = Kt.grad.double().cpu().numpy()
    Ga = oad.gp_nll_grad_K(K, y)
    rel = np.abs(G - Ga).max() / np.abs(Ga).max()
    print(f"FP32 GP gradient, relative error against the FP64 analytic gradient: {rel:.3e}")
    assert rel <= 2e-3, rel
