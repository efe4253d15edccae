"""GPU parity of the single-precision path and the mixed-precision solver:

* atlg_sgemm against an FP64 product of the same FP32 operands: every entry
  within K * eps32 * (|A||B|)_ij (ATLAS's gemmtst element bound, in FP32);
* atlg_sgemm_bf16 against an FP64 product of the operands rounded to
  bfloat16 in numpy (round to nearest even): BF16 x BF16 products are exact
  in FP32, so the same K * eps32 bound holds; the split form (hi + mid + lo
  operands, six products) against the same six products in FP64, and at
  FP32 accuracy against the FP32 product;
* atlg_sgetrf: lutst residual ||P A - L U||_1 / (||A||_1 eps n) <= 100 with
  eps of the arithmetic (FP32; BF16 updates: the bfloat16 unit roundoff),
  and pivots identical to the oracle's on matrices with well-separated
  pivots;
* atlg_dsgesv: LAPACK dsgesv semantics, and the normwise backward error
  ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) <= 1e-15 * n on
  A = U diag(s) V^T (U, V Haar orthogonal, s geometric from 1 to 1/kappa);
  converged solves leave A and B bit-identical; kappa beyond FP32's reach
  falls back to the FP64 factor (ITER < 0) and still meets the bound.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

EPS32 = np.finfo(np.float32).eps / 2
U_BF16 = 2.0 ** -8
U_BF16X3 = 2.0 ** -15  # two-term split operands: error of one product


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def dev(a):
    """Column-major device copy of a numpy matrix (as a flat tensor)."""
    import torch
    return torch.from_numpy(np.asfortranarray(a).reshape(-1, order="F").copy()).to("cuda")


def host(t, m, n, ld):
    return t.cpu().numpy().reshape(n, ld).T[:m, :].astype(np.float64)


def to_bf16(x):
    """float32 -> nearest-even bfloat16, returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(np.shape(x))


@pytest.mark.parametrize("ta,tb", [(111, 111), (112, 111), (111, 112), (112, 112)])
@pytest.mark.parametrize("m,n,k", [(257, 131, 333), (128, 128, 128), (1000, 700, 65)])
def test_sgemm(G, ta, tb, m, n, k):
    import torch
    rng = np.random.default_rng(m * 7 + n + k + ta + 3 * tb)
    a = rng.uniform(-1, 1, (k, m) if ta == 112 else (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (n, k) if tb == 112 else (k, n)).astype(np.float32)
    c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    da, db, dc = dev(a), dev(b), dev(c)
    G.sgemm(ta, tb, m, n, k, 1.5, da, a.shape[0], db, b.shape[0], -0.5, dc, m)
    torch.cuda.synchronize()
    got = host(dc, m, n, m)
    A = a.T if ta == 112 else a
    B = b.T if tb == 112 else b
    ref = 1.5 * (A.astype(np.float64) @ B.astype(np.float64)) - 0.5 * c.astype(np.float64)
    bound = (1.5 * np.abs(A).astype(np.float64) @ np.abs(B).astype(np.float64)
             + 0.5 * np.abs(c)) * (k + 2) * EPS32
    assert np.all(np.abs(got - ref) <= bound + 1e-30)


@pytest.mark.parametrize("split", [0, 1, 2])
@pytest.mark.parametrize("m,n,k", [(300, 200, 100), (128, 256, 64), (1031, 517, 777)])
def test_sgemm_bf16(G, m, n, k, split):
    import torch
    rng = np.random.default_rng(m + n + k)
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    c = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    da, db, dc = dev(a), dev(b), dev(c)
    G.sgemm_bf16(m, n, k, -1.0, da, m, db, k, 1.0, dc, m, split)
    torch.cuda.synchronize()
    got = host(dc, m, n, m)
    a_hi, b_hi = to_bf16(a), to_bf16(b)
    a16, b16 = a_hi.astype(np.float64), b_hi.astype(np.float64)
    if split == 1:
        a_mid = to_bf16(a - a_hi)
        b_mid = to_bf16(b - b_hi)
        a_lo = to_bf16(a - a_hi - a_mid).astype(np.float64)
        b_lo = to_bf16(b - b_hi - b_mid).astype(np.float64)
        a_mid, b_mid = a_mid.astype(np.float64), b_mid.astype(np.float64)
        terms = [(a16, b16), (a16, b_mid), (a_mid, b16), (a_mid, b_mid), (a16, b_lo), (a_lo, b16)]
        prod = sum(x @ y for x, y in terms)
        mag = sum(np.abs(x) @ np.abs(y) for x, y in terms)
    elif split == 2:
        a_lo = to_bf16(a - a_hi).astype(np.float64)
        b_lo = to_bf16(b - b_hi).astype(np.float64)
        terms = [(a16, b16), (a16, b_lo), (a_lo, b16)]
        prod = sum(x @ y for x, y in terms)
        mag = sum(np.abs(x) @ np.abs(y) for x, y in terms)
    else:
        prod = a16 @ b16
        mag = np.abs(a16) @ np.abs(b16)
    ref = c.astype(np.float64) - prod
    bound = (mag + np.abs(c)) * (6 * k + 2) * EPS32
    err = np.abs(got - ref)
    print("bf16 gemm split=%d %dx%dx%d: max err %.3g, max err/bound %.3g"
          % (split, m, n, k, err.max(), (err / (bound + 1e-30)).max()))
    assert np.all(err <= bound + 1e-30)
    ref32 = c.astype(np.float64) - a.astype(np.float64) @ b.astype(np.float64)
    err32 = np.abs(got - ref32).max()
    mag32 = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64)) + np.abs(c)
    if split == 1:
        # FP32 accuracy: within the FP32 GEMM element bound of the true product
        assert np.all(np.abs(got - ref32) <= 2 * mag32 * (k + 2) * EPS32)
    elif split == 2:
        # 16 significand bits per operand: about 2^-15 per product, far
        # below single BF16 and above FP32
        assert np.all(np.abs(got - ref32) <= mag32 * (2.0 ** -14 + (k + 2) * EPS32))
        # the lo terms were included: far closer than one BF16 term per operand
        assert err32 < 0.1 * np.abs(a16 @ b16 - a.astype(np.float64) @ b.astype(np.float64)).max()
    else:
        # the BF16 rounding really happened: an FP32 product differs
        assert err32 > 1e3 * err.max()


@pytest.mark.parametrize("split", [0, 1, 2])
@pytest.mark.parametrize("m,n,k", [(300, 200, 100), (1031, 517, 64)])
def test_sgemm_bf16_beta0(G, m, n, k, split):
    # beta = 0 overwrites C without reading it (NaN in C does not propagate);
    # m = 1031 takes the scalar C path (ldc not a multiple of 4)
    import torch
    rng = np.random.default_rng(m * 3 + k)
    a = rng.uniform(-1, 1, (m, k)).astype(np.float32)
    b = rng.uniform(-1, 1, (k, n)).astype(np.float32)
    da, db = dev(a), dev(b)
    dc = dev(np.full((m, n), np.nan, dtype=np.float32))
    G.sgemm_bf16(m, n, k, 2.0, da, m, db, k, 0.0, dc, m, split)
    torch.cuda.synchronize()
    got = host(dc, m, n, m)
    ref = 2.0 * (a.astype(np.float64) @ b.astype(np.float64))
    mag = 2.0 * (np.abs(a).astype(np.float64) @ np.abs(b).astype(np.float64))
    tol = mag * ({0: (k + 2) * 2.0 ** -8, 1: (6 * k + 2) * EPS32,
                  2: 2.0 ** -14 + (6 * k + 2) * EPS32}[split])
    assert np.all(np.isfinite(got))
    assert np.all(np.abs(got - ref) <= tol + 1e-30)


def lu_ratio(a0, lu, ipiv, eps):
    m, n = a0.shape
    mn = min(m, n)
    L = np.tril(lu[:, :mn], -1) + np.eye(m, mn)
    U = np.triu(lu[:mn, :])
    pa = a0.copy()
    for i in range(mn):
        p = ipiv[i]
        if p != i:
            pa[[i, p], :] = pa[[p, i], :]
    return np.abs(pa - L @ U).sum(axis=0).max() / (np.abs(a0).sum(axis=0).max() * eps * mn)


@pytest.mark.parametrize("prec", [0, 1, 2, 3])
@pytest.mark.parametrize("m,n", [(700, 700), (1536, 1536), (2000, 901), (901, 1300),
                                 (2304, 2304), (2600, 2100)])
def test_sgetrf(G, prec, m, n):
    # orders >= 2048 take the look-ahead driver, whose 512-deep updates use
    # the BF16 forms when prec != 0
    import torch
    rng = np.random.default_rng(m + n + prec)
    a = rng.uniform(-1, 1, (m, n)).astype(np.float32)
    da = dev(a)
    ipiv = torch.zeros(min(m, n), dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.sgetrf(m, n, da, m, ipiv, info, prec)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    lu = host(da, m, n, m)
    r = lu_ratio(a.astype(np.float64), lu, ipiv.cpu().numpy(),
                 {0: EPS32, 1: U_BF16, 2: EPS32, 3: U_BF16X3}[prec])
    print("sgetrf prec=%d %dx%d: ratio %.3g" % (prec, m, n, r))
    assert r <= 100.0


def test_sgetrf_pivots_match_oracle(G):
    """Well-separated pivots: the FP32 LU picks the oracle's rows."""
    import torch
    n = 640
    rng = np.random.default_rng(9)
    perm = rng.permutation(n)
    Lm = np.tril(rng.uniform(-1, 1, (n, n)) / n, -1) + np.eye(n)
    Um = np.triu(rng.uniform(-1, 1, (n, n)) / n, 1) + np.diag(rng.uniform(1, 2, n))
    a = (Lm @ Um)[np.argsort(perm)].astype(np.float32)
    da = dev(a)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.sgetrf(n, n, da, n, ipiv, info, 0)
    torch.cuda.synchronize()
    ref = np.asfortranarray(a.astype(np.float64))
    rp, rinfo = O.getf2(ref)
    assert rinfo == 0
    assert np.array_equal(ipiv.cpu().numpy(), rp)


def test_sgetrs(G):
    import torch
    n, nrhs = 900, 3
    rng = np.random.default_rng(12)
    a = rng.uniform(-1, 1, (n, n)).astype(np.float32) + 4 * np.eye(n, dtype=np.float32)
    b = rng.uniform(-1, 1, (n, nrhs)).astype(np.float32)
    da, db = dev(a), dev(b)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.sgetrf(n, n, da, n, ipiv, info, 0)
    G.sgetrs(111, n, nrhs, da, n, ipiv, db, n)
    torch.cuda.synchronize()
    x = host(db, n, nrhs, n)
    A = a.astype(np.float64)
    r = np.abs(b - A @ x).max(axis=0)
    be = r / (np.abs(A).sum(axis=1).max() * np.abs(x).max(axis=0) * n * EPS32)
    assert be.max() <= 100.0


def haar(n, rng):
    q, r = np.linalg.qr(rng.standard_normal((n, n)))
    return q * np.sign(np.diag(r))


def cond_matrix(n, kappa, rng):
    u, v = haar(n, rng), haar(n, rng)
    s = kappa ** (-np.arange(n) / (n - 1))
    return (u * s) @ v.T


def backward_error(a, x, b):
    r = np.abs(b - a @ x).max(axis=0)
    return (r / (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0)
                 + np.abs(b).max(axis=0))).max()


@pytest.mark.parametrize("lowprec", [0, 1, 2, 3])
@pytest.mark.parametrize("n,nrhs,kappa", [(1024, 1, 1e2), (1536, 1, 1e6), (777, 3, 1e2),
                                          (1024, 1, 1e10), (600, 6, 1e6)])
def test_dsgesv(G, lowprec, n, nrhs, kappa):
    import torch
    rng = np.random.default_rng(int(n + nrhs + np.log10(kappa)))
    a = cond_matrix(n, kappa, rng)
    b = rng.standard_normal((n, nrhs))
    da, db = dev(a), dev(b)
    dx = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    it = G.dsgesv(n, nrhs, da, n, ipiv, db, n, dx, n, info, 30, lowprec)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    x = host(dx, n, nrhs, n)
    be = backward_error(a, x, b)
    print("dsgesv lowprec=%d n=%d nrhs=%d kappa=%g: ITER=%d backward error %.3g"
          % (lowprec, n, nrhs, kappa, it, be))
    assert be <= 1e-15 * n
    assert np.array_equal(host(db, n, nrhs, n), b)  # B is input only
    if it >= 0:
        assert np.array_equal(host(da, n, n, n), a)  # A untouched after refinement
    else:
        assert it in (-2, -3, -31)
        lu = host(da, n, n, n)
        assert lu_ratio(a, lu, ipiv.cpu().numpy(), np.finfo(np.float64).eps / 2) <= 100.0
    if kappa <= 1e6 and lowprec in (0, 2):
        assert it >= 0  # FP32-accurate factor: refinement converges for kappa * u32 < 1
    if kappa <= 1e2 and lowprec == 3:
        assert it >= 0  # two-term split factor: kappa * 2^-15 well below 1
    if kappa >= 1e10 and lowprec in (0, 2, 3):
        assert it < 0


@pytest.mark.parametrize("lowprec", [0, 1, 2, 3])
@pytest.mark.parametrize("n,nrhs,kappa", [(1024, 1, 1e2), (1536, 1, 1e6), (600, 2, 1e6),
                                          (1024, 1, 1e10)])
def test_dsgesv_gmres_ir(G, lowprec, n, nrhs, kappa):
    """GMRES-IR (refine=1): the low-precision factor preconditions FP64
    GMRES; every solve meets the FP64 backward-error bound (by refinement or
    the FP64 fallback), converged ones with A and B untouched."""
    import torch
    rng = np.random.default_rng(int(7 * n + nrhs + np.log10(kappa)))
    a = cond_matrix(n, kappa, rng)
    b = rng.standard_normal((n, nrhs))
    da, db = dev(a), dev(b)
    dx = torch.zeros(n * nrhs, dtype=torch.float64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    it, inner = G.dsgesv(n, nrhs, da, n, ipiv, db, n, dx, n, info, 30, lowprec, refine=1,
                         inner=True)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    x = host(dx, n, nrhs, n)
    be = backward_error(a, x, b)
    print("gmres-ir lowprec=%d n=%d nrhs=%d kappa=%g: ITER=%d inner=%d backward error %.3g"
          % (lowprec, n, nrhs, kappa, it, inner, be))
    assert be <= 1e-15 * n
    assert np.array_equal(host(db, n, nrhs, n), b)
    if it >= 0:
        assert np.array_equal(host(da, n, n, n), a)
    # GMRES-IR converges while kappa * u_factor stays well below 1/u_factor:
    # FP32-accurate factors (FP32, split BF16) to kappa = 1e6; the BF16
    # factor (u = 2^-8, beyond the reach of classical refinement already at
    # kappa = 1e2 for larger n) and the two-term split one at kappa = 1e2
    if kappa <= (1e6 if lowprec in (0, 2) else 1e2):
        assert it >= 0


@pytest.mark.parametrize("lowprec", [0, 2])
def test_dsgesv_gmres_ir_n16384_kappa1e6(G, lowprec):
    """GMRES-IR at n = 16384, kappa = 1e6 with the FP32-accurate factors
    (FP32, split BF16): it must converge without the FP64 fallback and meet
    the backward-error bound ||b - A x|| / (||A|| ||x|| + ||b||) <= 1e-15 n.
    A = U diag(s) V^T is generated on the GPU (Haar U, V; s geometric from
    1 to 1e-6); the residual is an FP64 torch product on the device."""
    import torch
    n = 16384
    g = torch.Generator(device="cuda")
    g.manual_seed(16384 + lowprec)

    def haar():
        q, r = torch.linalg.qr(torch.randn((n, n), generator=g, dtype=torch.float64,
                                           device="cuda"))
        return q * torch.sign(torch.diagonal(r)).unsqueeze(0)

    U, V = haar(), haar()
    s = torch.pow(torch.tensor(1e6, dtype=torch.float64, device="cuda"),
                  -torch.arange(n, dtype=torch.float64, device="cuda") / (n - 1))
    At = torch.mm(V * s.unsqueeze(0), U.t().contiguous())  # row-major A^T = column-major A
    del U, V
    A = At.clone()
    b = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    it, inner = G.dsgesv(n, 1, A, n, ipiv, b, n, x, n, info, 30, lowprec, refine=1, inner=True)
    torch.cuda.synchronize()
    r = b - torch.mv(At.t(), x)
    anrm = float(At.abs().sum(dim=0).max())
    be = float(r.abs().max()) / (anrm * float(x.abs().max()) + float(b.abs().max()))
    print("gmres-ir n=16384 kappa=1e6 lowprec=%d: ITER=%d inner=%d backward error %.3g"
          % (lowprec, it, inner, be))
    assert int(info.item()) == 0
    assert it >= 0, (it, inner)
    assert be <= 1e-15 * n
    assert torch.equal(A, At)  # converged: A untouched
