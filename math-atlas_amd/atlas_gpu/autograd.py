"""Differentiable Cholesky path over the C ABI (torch.autograd.Function).

Forward and backward both run in libatlas_gpu.so: potrf, the triangular
solve and logdet forward, and the hand-written adjoint kernels
(atlg_dpotrf_adjoint, atlg_dtrsm_adjoint, atlg_dlogdet_chol_adjoint)
backward of the primitives.  The fused log-likelihood's backward uses the
closed form of its gradient (potri, a triangular solve and a SYRK).  Matrices are torch float64 CUDA tensors in their natural
orientation (K symmetric, L lower triangular with K = L L^T); the column-major
buffers the library works on are built internally.

Primitives, composable under autograd:
    L = cholesky(K)            lower factor
    a = solve_lower(L, y)      a = L^-1 y  (y a vector or N x k)
    d = logdet_chol(L)         log det K = 2 sum log L_ii

Fused objective with a memory policy:
    gaussian_loglik(K, y, store_factor=None)
        = -1/2 y^T K^-1 y - 1/2 log det K - N/2 log(2 pi)
    (for y with k columns: the sum of the k columns' log-likelihoods)
    store_factor=True keeps L and a for the backward pass (N^2 + N doubles);
    False keeps only the inputs and recomputes the factor in backward;
    None decides from free device memory.
"""
import math

import torch

from . import (Lower, NoTrans, Trans, NonUnit, Left, dpotrf, dpotri, dsyrk, dtrsm,
               dpotrf_adjoint, dtrsm_adjoint, dlogdet_chol, dlogdet_chol_adjoint)


def _check_info(info, what):
    v = int(info.item())
    if v != 0:
        raise torch.linalg.LinAlgError("%s: leading minor %d is not positive definite" % (what, v))


def _factor(K):
    """Column-major lower Cholesky buffer of symmetric K (upper part: K)."""
    n = K.shape[0]
    buf = K.detach().to(torch.float64).contiguous().clone()
    info = torch.zeros(1, dtype=torch.int32, device=K.device)
    dpotrf(Lower, n, buf, n, info)
    _check_info(info, "potrf")
    return buf


def _cm(t):
    """Column-major buffer of a 2-D tensor (element (i,j) at j*N + i)."""
    return t.t().contiguous()


def _rhs_buf(y):
    """Column-major N x k buffer of y (vector or matrix), and k."""
    if y.dim() == 1:
        return y.detach().contiguous().clone(), 1
    return _cm(y.detach()), y.shape[1]


def _rhs_out(buf, like):
    return buf if like.dim() == 1 else buf.t()


class Potrf(torch.autograd.Function):
    @staticmethod
    def forward(ctx, K):
        buf = _factor(K)
        L = torch.tril(buf.t())
        ctx.save_for_backward(L)
        return L

    @staticmethod
    def backward(ctx, Lbar):
        (L,) = ctx.saved_tensors
        n = L.shape[0]
        Lc = _cm(L)
        G = _cm(Lbar.to(torch.float64))
        dpotrf_adjoint(n, Lc, n, G, n)
        return G  # symmetric, so the column-major buffer is the tensor


class SolveLower(torch.autograd.Function):
    @staticmethod
    def forward(ctx, L, y):
        n = L.shape[0]
        Lc = _cm(L.detach())
        X, k = _rhs_buf(y)
        dtrsm(Left, Lower, NoTrans, NonUnit, n, k, 1.0, Lc, n, X, n)
        ctx.save_for_backward(L, X)
        ctx.k, ctx.vec = k, y.dim() == 1
        return _rhs_out(X, y).clone()

    @staticmethod
    def backward(ctx, abar):
        L, X = ctx.saved_tensors
        n, k = L.shape[0], ctx.k
        Lc = _cm(L)
        Xbar = abar.detach().contiguous().clone() if ctx.vec else _cm(abar.detach())
        Lbar = torch.zeros((n, n), dtype=torch.float64, device=L.device)
        dtrsm_adjoint(n, k, Lc, n, X, n, Xbar, n, Lbar, n)
        ybar = Xbar if ctx.vec else Xbar.t()
        return Lbar.t(), ybar


class LogdetChol(torch.autograd.Function):
    @staticmethod
    def forward(ctx, L):
        n = L.shape[0]
        Lc = L.detach().contiguous()  # the diagonal is the same in either orientation
        out = torch.zeros(1, dtype=torch.float64, device=L.device)
        dlogdet_chol(n, Lc, n, out)
        ctx.save_for_backward(Lc)
        return out[0]

    @staticmethod
    def backward(ctx, dbar):
        (Lc,) = ctx.saved_tensors
        n = Lc.shape[0]
        Lbar = torch.zeros((n, n), dtype=torch.float64, device=Lc.device)
        g = dbar.detach().reshape(1).to(torch.float64).contiguous()
        dlogdet_chol_adjoint(n, Lc, n, g, Lbar, n)
        return Lbar


def cholesky(K):
    return Potrf.apply(K)


def solve_lower(L, y):
    return SolveLower.apply(L, y)


def logdet_chol(L):
    return LogdetChol.apply(L)


def _store_default(n, device):
    free, _ = torch.cuda.mem_get_info(device)
    return 8.0 * n * n < 0.25 * free


class GaussianLogLik(torch.autograd.Function):
    """-1/2 y^T K^-1 y - 1/2 log det K - N/2 log 2pi, one fused node."""

    @staticmethod
    def forward(ctx, K, y, store_factor):
        n = K.shape[0]
        buf = _factor(K)
        X, k = _rhs_buf(y)
        dtrsm(Left, Lower, NoTrans, NonUnit, n, k, 1.0, buf, n, X, n)
        ld = torch.zeros(1, dtype=torch.float64, device=K.device)
        dlogdet_chol(n, buf, n, ld)
        val = -0.5 * (X * X).sum() - 0.5 * k * ld[0] - 0.5 * n * k * math.log(2.0 * math.pi)
        ctx.store = bool(_store_default(n, K.device) if store_factor is None else store_factor)
        ctx.k, ctx.vec = k, y.dim() == 1
        if ctx.store:
            ctx.save_for_backward(K, y, buf, X)
        else:
            ctx.save_for_backward(K, y)
        return val

    @staticmethod
    def backward(ctx, g):
        n, k = ctx.saved_tensors[0].shape[0], ctx.k
        if ctx.store:
            K, y, buf, X = ctx.saved_tensors
        else:
            K, y = ctx.saved_tensors
            buf = _factor(K)  # recompute the factor and a = L^-1 y
            X, _ = _rhs_buf(y)
            dtrsm(Left, Lower, NoTrans, NonUnit, n, k, 1.0, buf, n, X, n)
        gval = g.detach().to(torch.float64)
        # The node's gradient in closed form (Rasmussen & Williams 2006,
        # eq. 5.9): with A = K^-1 Y = L^-T X,
        #   dloss/dK = g/2 (A A^T - k K^-1),   dloss/dY = -g A,
        # K^-1 = L^-T L^-1 by potri (trtri + lauum, 2N^3/3 flops) in place of
        # the chain of the three adjoints (about 3N^3).  The lower triangle
        # of K^-1 takes the rank-k term by SYRK, then is mirrored.
        A = X.clone()
        dtrsm(Left, Lower, Trans, NonUnit, n, k, 1.0, buf, n, A, n)
        W = buf.clone() if ctx.store else buf
        info = torch.zeros(1, dtype=torch.int32, device=W.device)
        dpotri(Lower, n, W, n, info)  # cannot fail: the factor's diagonal is positive
        dsyrk(Lower, NoTrans, n, k, 0.5, A, n, -0.5 * k, W, n)
        # W's column-major lower triangle is the tensor's upper one
        G = torch.triu(W).add_(torch.triu(W, 1).t()).mul_(gval)
        ybar = _rhs_out(A.mul_(-gval), y)
        return G, ybar, None


def gaussian_loglik(K, y, store_factor=None):
    return GaussianLogLik.apply(K, y, store_factor)
