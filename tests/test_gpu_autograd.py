"""Differentiable Cholesky log-likelihood on the GPU (adjoint kernels of
potrf, trsm and logdet behind torch.autograd.Function), checked against
* the CPU FP64 torch autograd gradient at N=1024 (relative error <= 1e-9),
* central finite differences, h = 1e-6, on 20 random directions
  (symmetric in K, arbitrary in y) at N=1024 and at the full N=8192
  (relative error of the directional derivative <= 1e-4; the finite
  difference itself carries ~eps*|l|/h of rounding error).
Workload: K = RBF kernel, length-scale 0.2, over N points iid U[0,1]^3, plus
1e-2 I; y iid N(0,1)."""
import math

import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def AG():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    from atlas_gpu import autograd
    return autograd


def rbf_problem(n, seed, device="cuda"):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    x = torch.rand((n, 3), generator=g, dtype=torch.float64)
    y = torch.randn((n,), generator=g, dtype=torch.float64)
    d2 = ((x[:, None, :] - x[None, :, :]) ** 2).sum(-1) if n <= 2048 else None
    if d2 is None:
        xd = x.to(device)
        d2 = (xd * xd).sum(1)[:, None] + (xd * xd).sum(1)[None, :] - 2.0 * (xd @ xd.t())
        d2 = d2.clamp_min(0.0)
        K = torch.exp(-d2 / (2 * 0.2 ** 2))
        K = 0.5 * (K + K.t())
        K.diagonal().add_(1e-2)
        return K, y.to(device)
    K = torch.exp(-d2 / (2 * 0.2 ** 2))
    K.diagonal().add_(1e-2)
    return K.to(device), y.to(device)


def cpu_reference(K, y):
    import torch
    Kc = K.detach().cpu().clone().requires_grad_(True)
    yc = y.detach().cpu().clone().requires_grad_(True)
    L = torch.linalg.cholesky(Kc)
    a = torch.linalg.solve_triangular(L, yc[:, None], upper=False)[:, 0]
    n = K.shape[0]
    val = -0.5 * (a * a).sum() - torch.log(torch.diagonal(L)).sum() - 0.5 * n * math.log(2 * math.pi)
    val.backward()
    gK = 0.5 * (Kc.grad + Kc.grad.t())
    return val.item(), gK, yc.grad


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.mark.parametrize("store", [True, False])
def test_loglik_matches_cpu_autograd_1024(AG, store):
    import torch
    K, y = rbf_problem(1024, 1)
    v_ref, gK_ref, gy_ref = cpu_reference(K, y)
    Kg = K.clone().requires_grad_(True)
    yg = y.clone().requires_grad_(True)
    val = AG.gaussian_loglik(Kg, yg, store_factor=store)
    val.backward()
    assert abs(val.item() - v_ref) <= 1e-10 * abs(v_ref)
    eK, ey = rel(Kg.grad.cpu(), gK_ref), rel(yg.grad.cpu(), gy_ref)
    print("N=1024 store=%s: value rel %.2e, dK rel %.2e, dy rel %.2e" % (
        store, abs(val.item() - v_ref) / abs(v_ref), eK, ey))
    assert eK <= 1e-9 and ey <= 1e-9


def test_primitives_compose_to_same_gradient(AG):
    import torch
    K, y = rbf_problem(600, 2)
    Kg = K.clone().requires_grad_(True)
    yg = y.clone().requires_grad_(True)
    L = AG.cholesky(Kg)
    a = AG.solve_lower(L, yg)
    val = -0.5 * (a * a).sum() - 0.5 * AG.logdet_chol(L) - 0.5 * 600 * math.log(2 * math.pi)
    val.backward()
    v_ref, gK_ref, gy_ref = cpu_reference(K, y)
    assert abs(val.item() - v_ref) <= 1e-10 * abs(v_ref)
    assert rel(Kg.grad.cpu(), gK_ref) <= 1e-9
    assert rel(yg.grad.cpu(), gy_ref) <= 1e-9


@pytest.mark.parametrize("n", [1024, 8192])
def test_loglik_finite_differences(AG, n):
    import torch
    K, y = rbf_problem(n, 3)
    Kg = K.clone().requires_grad_(True)
    yg = y.clone().requires_grad_(True)
    val = AG.gaussian_loglik(Kg, yg, store_factor=(n <= 1024))
    val.backward()
    gK, gy = Kg.grad.detach(), yg.grad.detach()
    gen = torch.Generator(device="cuda").manual_seed(4)
    h = 1e-6
    worst = 0.0
    for _ in range(20):
        D = torch.randn((n, n), generator=gen, dtype=torch.float64, device="cuda")
        D = 0.5 * (D + D.t())
        D /= torch.linalg.norm(D)
        dy = torch.randn((n,), generator=gen, dtype=torch.float64, device="cuda")
        dy /= torch.linalg.norm(dy)
        with torch.no_grad():
            fp = AG.gaussian_loglik(K + h * D, y + h * dy, store_factor=False).item()
            fm = AG.gaussian_loglik(K - h * D, y - h * dy, store_factor=False).item()
        fd = (fp - fm) / (2 * h)
        an = float((gK * D).sum() + (gy * dy).sum())
        err = abs(fd - an) / max(abs(an), 1e-300)
        worst = max(worst, err)
        # FD accuracy: truncation O(h^2) plus rounding, each evaluation of
        # the log-likelihood being accurate to about u * n * |f|, divided by h
        noise = 2.0 * 1.1e-16 * n * abs(val.item()) / (h * max(abs(an), 1e-300))
        assert err <= 1e-4 + noise, (err, noise)
    print("N=%d: worst FD relative error over 20 directions %.2e (loglik %.6e)" % (n, worst, val.item()))


def test_loglik_several_columns_matches_cpu_autograd(AG):
    # y with k = 3 columns: the sum of the columns' log-likelihoods, the
    # gradient in closed form against CPU FP64 autograd through
    # torch.linalg.cholesky and a triangular solve
    import torch
    n, k = 700, 3
    K, _ = rbf_problem(n, 6)
    g = torch.Generator().manual_seed(6)
    Y = torch.randn((n, k), generator=g, dtype=torch.float64).to(K.device)
    Kc = K.detach().cpu().clone().requires_grad_(True)
    Yc = Y.detach().cpu().clone().requires_grad_(True)
    L = torch.linalg.cholesky(Kc)
    a = torch.linalg.solve_triangular(L, Yc, upper=False)
    ref = -0.5 * (a * a).sum() - k * torch.log(torch.diagonal(L)).sum() \
        - 0.5 * n * k * math.log(2 * math.pi)
    (2.0 * ref).backward()  # an upstream gradient other than 1
    gK_ref = 0.5 * (Kc.grad + Kc.grad.t())
    for store in (True, False):
        Kg = K.clone().requires_grad_(True)
        Yg = Y.clone().requires_grad_(True)
        val = AG.gaussian_loglik(Kg, Yg, store_factor=store)
        (2.0 * val).backward()
        assert abs(val.item() - ref.item()) <= 1e-10 * abs(ref.item())
        assert rel(Kg.grad.cpu(), gK_ref) <= 1e-9
        assert rel(Yg.grad.cpu(), Yc.grad) <= 1e-9
        assert torch.equal(Kg.grad, Kg.grad.t())


@pytest.mark.parametrize("n", [300, 1500])
def test_potrf_adjoint_matches_closed_form(AG, n):
    # the primitives' chain (potrf adjoint with trimmed dense products, the
    # few-RHS solve adjoint, logdet) against the fused node's closed-form
    # gradient on the same problem, at orders with ragged 128-blocks
    import torch
    K, y = rbf_problem(n, 8)
    Kg = K.clone().requires_grad_(True)
    yg = y.clone().requires_grad_(True)
    L = AG.cholesky(Kg)
    a = AG.solve_lower(L, yg)
    val = -0.5 * (a * a).sum() - 0.5 * AG.logdet_chol(L) - 0.5 * n * math.log(2 * math.pi)
    val.backward()
    Kf = K.clone().requires_grad_(True)
    yf = y.clone().requires_grad_(True)
    AG.gaussian_loglik(Kf, yf, store_factor=True).backward()
    assert rel(Kg.grad, Kf.grad) <= 1e-9
    assert rel(yg.grad, yf.grad) <= 1e-9
