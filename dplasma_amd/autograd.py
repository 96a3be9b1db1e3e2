"""Differentiable Cholesky for PyTorch over the C ABI of libdplasma_mi.so.

forward:  mi_dpotrf_tiles / mi_spotrf_tiles (the tile factorization),
backward: mi_dpotrf_adjoint_tiles / mi_spotrf_adjoint_tiles (its blocked
          reverse), returning the symmetric gradient with respect to A.

A thin wrapper: dense N x N device tensors are packed into the lower tile
layout of the kernels (tile (m, n), m >= n, column by column, each NB x NB
column-major) and back with torch indexing; the kernels run on the library's
own streams, so the wrapper synchronizes the device around each call.  With
recompute=True only A is kept for the backward pass and L is factored again
there (for when memory is short); otherwise the factor's tiles are kept.
"""
import ctypes
import math

import torch

from . import Device


class _Layout:
    """Index tensors of the packed lower tile order for one (N, NB, device)."""

    _cache = {}

    @classmethod
    def get(cls, N, NB, device):
        key = (N, NB, str(device))
        if key not in cls._cache:
            MT = (N + NB - 1) // NB
            ns = [n for n in range(MT) for m in range(n, MT)]
            ms = [m for n in range(MT) for m in range(n, MT)]
            cls._cache[key] = (MT, torch.tensor(ns, device=device), torch.tensor(ms, device=device))
        return cls._cache[key]


def to_tiles(A, NB):
    """Dense (N x N) -> packed lower tiles (ntiles, NB, NB), tile [column][row]."""
    N = A.shape[0]
    MT, ns, ms = _Layout.get(N, NB, A.device)
    P = MT * NB
    if P != N:
        Ap = A.new_zeros(P, P)
        Ap[:N, :N] = A
    else:
        Ap = A.contiguous()
    T4 = Ap.view(MT, NB, MT, NB).permute(2, 0, 3, 1)     # [n][m][column][row]
    return T4[ns, ms].contiguous()


def from_tiles(T, N, NB):
    """Packed lower tiles -> dense lower triangle (N x N)."""
    MT, ns, ms = _Layout.get(N, NB, T.device)
    T4 = T.new_zeros(MT, MT, NB, NB)
    T4[ns, ms] = T
    return T4.permute(1, 3, 0, 2).reshape(MT * NB, MT * NB)[:N, :N].tril()


class _Lib:
    dev = None

    @classmethod
    def get(cls):
        if cls.dev is None:
            cls.dev = Device(torch.cuda.current_device())
        return cls.dev


def _potrf(T, N, NB):
    d = _Lib.get()
    info = ctypes.c_int(-1)
    fn = d.M.mi_dpotrf_tiles if T.dtype == torch.float64 else d.M.mi_spotrf_tiles
    torch.cuda.synchronize()
    d.check(fn(d.h, N, NB, ctypes.c_void_p(T.data_ptr()), ctypes.byref(info)), "potrf")
    if info.value != 0:
        raise ValueError(f"matrix is not positive definite (info = {info.value})")


def _adjoint(T, G, N, NB):
    d = _Lib.get()
    fn = d.M.mi_dpotrf_adjoint_tiles if T.dtype == torch.float64 else d.M.mi_spotrf_adjoint_tiles
    torch.cuda.synchronize()
    d.check(fn(d.h, N, NB, ctypes.c_void_p(T.data_ptr()), ctypes.c_void_p(G.data_ptr())), "potrf adjoint")


class Cholesky(torch.autograd.Function):
    """L = cholesky(A) (lower) on the tile kernels; backward on the adjoint kernels."""

    @staticmethod
    def forward(ctx, A, NB=512, recompute=False):
        if A.dtype not in (torch.float64, torch.float32) or A.device.type != "cuda":
            raise TypeError("Cholesky: A must be a float64 or float32 CUDA tensor")
        N = A.shape[0]
        T = to_tiles(A.detach(), NB)
        _potrf(T, N, NB)
        ctx.NB, ctx.N, ctx.recompute = NB, N, recompute
        ctx.save_for_backward(A.detach() if recompute else T)
        return from_tiles(T, N, NB)

    @staticmethod
    def backward(ctx, gL):
        (saved,) = ctx.saved_tensors
        N, NB = ctx.N, ctx.NB
        if ctx.recompute:
            T = to_tiles(saved, NB)
            _potrf(T, N, NB)
        else:
            T = saved
        G = to_tiles(gL.tril().to(T.dtype), NB)
        _adjoint(T, G, N, NB)
        Ab = from_tiles(G, N, NB)
        return Ab + Ab.tril(-1).T, None, None


def cholesky(A, NB=512, recompute=False):
    return Cholesky.apply(A, NB, recompute)


def se_kernel(X, lengthscale, variance=1.0, noise=1e-2):
    """Squared-exponential kernel v exp(-|x_i - x_j|^2 / (2 l^2)) + noise I."""
    d2 = (X * X).sum(1)[:, None] + (X * X).sum(1)[None, :] - 2.0 * X @ X.T
    d2 = d2.clamp_min(0.0)
    K = variance * torch.exp(-0.5 * d2 / lengthscale ** 2)
    return K + noise * torch.eye(X.shape[0], dtype=X.dtype, device=X.device)


def gp_nll(K, y, NB=512, recompute=False):
    """Negative log marginal likelihood of a Gaussian process:
    1/2 y^T K^-1 y + 1/2 logdet K + N/2 log 2 pi, with K = L L^T from the tile
    kernels: two triangular solves and logdet = 2 sum log diag(L)."""
    L = cholesky(K, NB, recompute)
    a = torch.linalg.solve_triangular(L, y.reshape(-1, 1), upper=False)
    return 0.5 * (a * a).sum() + torch.log(torch.diagonal(L)).sum() + 0.5 * y.numel() * math.log(2 * math.pi)
