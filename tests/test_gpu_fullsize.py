"""Parity at the full sizes of the level-3 sweep (one MI355X).

At these orders the CPU oracle cannot redo the whole operation in seconds,
so each case combines
* the oracle on a sample: the textbook GEMM (ATL_drefgemm loop order) on a
  few rows of C, the textbook TRSM on a few right-hand sides, judged with
  l3blastst's residual ratio (< 50) and gemmtst's element-wise bound
  (AtlasBase/Antoine/atlbinblas.base:11791-11935, 12761-12900;
  AtlasBase/Clint/atlas.base:21968);
* size-independent checks of the whole result computed in FP64 on the GPU
  with torch (the checker, not the product): Freivalds' test
  C x = A (B x) + beta C0 x for GEMM, the TRSM residual A X - B, and
  llttst's ||A - L L^T||_1 / (||A||_1 eps N) <= 16 for Cholesky
  (AtlasBase/Clint/atlas-tlp.base:1320-1514), plus a known factor
  A = L0 L0^T (the factor with positive diagonal is unique), the unused
  triangle left untouched, and info from a minor in a late block.
Matrices are column-major: a torch (cols, ld) tensor t holds t^T.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def _uniform(torch, n, m, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.empty((n, m), dtype=torch.float64, device="cuda").uniform_(-1.0, 1.0, generator=g)


@pytest.mark.parametrize("n", [4096, 8192, 16384])
def test_dgemm_square_full(G, n):
    import torch
    a, b, c0 = (_uniform(torch, n, n, 100 * n + i) for i in range(3))
    c = c0.clone()
    G.dgemm(G.NoTrans, G.NoTrans, n, n, n, 1.0, a, n, b, n, 1.0, c, n)
    torch.cuda.synchronize()
    # oracle on sampled rows of C (rows of the column-major A are columns of a)
    rows = torch.tensor([0, 1, n // 3, n // 2 + 7, n - 129, n - 1], device="cuda")
    ar = np.asfortranarray(a[:, rows].t().cpu().numpy())          # r x n
    bm = G.to_fortran(b)                                          # n x n
    c0r = np.asfortranarray(c0[:, rows].t().cpu().numpy())        # r x n
    got = np.asfortranarray(c[:, rows].t().cpu().numpy())
    ref = c0r.copy(order="F")
    r = len(rows)
    O.refgemm(O.NoTrans, O.NoTrans, r, n, n, 1.0, ar, bm, 1.0, ref)
    ratio = O.gemm_resid(O.NoTrans, O.NoTrans, r, n, n, 1.0, ar, bm, 1.0, c0r, got, ref)
    nbad, mx = O.gemm_elem_check(r, n, n, 1.0, 1.0, got, ref)
    print("dgemm n=%d: l3blastst ratio %.3g, max elem diff %.3g" % (n, ratio, mx))
    assert ratio < 50.0 and nbad == 0
    del bm
    # Freivalds on the whole product: C x = A (B x) + C0 x within 2 n eps of
    # (|A| (|B| |x|) + |C0| |x|), elementwise
    g = torch.Generator(device="cuda")
    g.manual_seed(n + 5)
    x = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    lhs = torch.mv(c.t(), x)
    rhs = torch.mv(a.t(), torch.mv(b.t(), x)) + torch.mv(c0.t(), x)
    bound = torch.mv(a.t().abs(), torch.mv(b.t().abs(), x.abs())) + torch.mv(c0.t().abs(), x.abs())
    q = float(((lhs - rhs).abs() / bound).max()) / (n * 2.0 ** -52)
    print("dgemm n=%d: Freivalds |Cx - ABx - C0x| / (n eps bound) = %.3g" % (n, q))
    assert q <= 2.0


def test_dtrsm_llnu_full(G):
    import torch
    n = 16384
    a = _uniform(torch, n, n, 30).mul_(1.0 / n)  # strict lower scaled by 1/n; unit diagonal unread
    b0 = _uniform(torch, n, n, 31)
    b = b0.clone()
    G.dtrsm(G.Left, G.Lower, G.NoTrans, G.Unit, n, n, 1.0, a, n, b, n)
    torch.cuda.synchronize()
    # oracle on sampled right-hand sides
    cols = torch.tensor([0, 5000, n - 1], device="cuda")
    am = G.to_fortran(a)
    b0c = np.asfortranarray(b0[cols].t().cpu().numpy())
    got = np.asfortranarray(b[cols].t().cpu().numpy())
    ref = b0c.copy(order="F")
    O.reftrsm(O.Left, O.Lower, O.NoTrans, O.Unit, n, len(cols), 1.0, am, ref)
    ratio = O.trsm_resid(O.Left, O.Lower, O.Unit, n, len(cols), 1.0, am, b0c, got, ref)
    print("dtrsm 16384: l3blastst ratio %.3g on sampled columns" % ratio)
    assert ratio < 50.0
    del am
    # every right-hand side: ||L X - B||_1 / (||L||_1 ||X||_1 n eps)
    low = torch.tril(a.t(), -1)
    low.diagonal().fill_(1.0)
    res = torch.matmul(low, b.t()).sub_(b0.t())
    q = float(res.abs().sum(dim=0).max()) / (
        float(low.abs().sum(dim=0).max()) * float(b.t().abs().sum(dim=0).max()) * n * 2.0 ** -52)
    print("dtrsm 16384: full residual ratio %.3g" % q)
    assert q <= 1.0


def _spd(torch, n, seed):
    # B B^T / n + n I with B iid normal(0, 1): the sweep's SPD matrices
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    bm = torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda")
    s = torch.mm(bm, bm.t()).div_(n)
    s = (s + s.t()).mul_(0.5)
    s.diagonal().add_(float(n))
    return s


def _unused_mask(torch, n, uplo_lower):
    # column-major Lower leaves i < j unused: tensor entries t[j, i], j > i
    return torch.ones((n, n), dtype=torch.bool, device="cuda").tril_(-1) if uplo_lower else \
        torch.ones((n, n), dtype=torch.bool, device="cuda").triu_(1)


@pytest.mark.parametrize("n,lower", [(12288, True), (12288, False), (16384, True), (16384, False)])
def test_dpotrf_full_llttst(G, n, lower):
    # 12288 takes the 256-column look-ahead steps, 16384 the 384-column ones
    import torch
    s = _spd(torch, n, 7 * n + lower)
    f = s.clone()
    mask = _unused_mask(torch, n, lower)
    f[mask] = -973200000.0
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.dpotrf(G.Lower if lower else G.Upper, n, f, n, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert bool((f[mask] == -973200000.0).all())
    fm = f.t()
    if lower:
        l = torch.tril(fm)
        prod = torch.matmul(l, l.t().contiguous())
    else:
        u = torch.triu(fm)
        prod = torch.matmul(u.t().contiguous(), u)
    del f, fm
    prod.sub_(s)
    r = float(prod.abs().sum(dim=0).max()) / (float(s.abs().sum(dim=0).max()) * O.epsilon() * n)
    print("dpotrf n=%d %s: llttst ratio %.3g" % (n, "L" if lower else "U", r))
    assert r <= 16.0


def test_dpotrf_full_known_factor_and_info(G):
    import torch
    n = 16384
    g = torch.Generator(device="cuda")
    g.manual_seed(11)
    l0 = torch.empty((n, n), dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=g)
    l0 = torch.tril(l0, -1).div_(n)
    l0.diagonal().copy_(torch.rand(n, generator=g, dtype=torch.float64, device="cuda") + 1.0)
    a = torch.matmul(l0, l0.t().contiguous())   # row-major A = column-major A (symmetric)
    a = (a + a.t()).mul_(0.5)
    f = a.clone()
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.dpotrf(G.Lower, n, f, n, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    l = torch.tril(f.t())
    err = float((l - l0).abs().max())
    print("dpotrf 16384 known factor: max |L - L0| = %.3g" % err)
    assert err <= 1e-10
    del l, f
    # a non-positive minor in a late block: info = its order, as ATLAS stops
    k = 14000
    a[k, k] = -1.0e6
    info.zero_()
    G.dpotrf(G.Lower, n, a, n, info)
    torch.cuda.synchronize()
    assert int(info.item()) == k + 1
