"""GPU parity of the batched SPD solvers (atlg_dpotrf/dpotrs/dposv_batched,
the FP32 forms and the FP32-factor + FP64-refinement atlg_dsposv_batched)
against the CPU oracle, system by system, with ATLAS's tester formulas:

* llttst: ||A - L L^T||_1 / (||A||_1 eps n) <= 16 (AtlasTest's -T 16,
  AtlasBase/Clint/atlas-tlp.base:1320-1514), with eps of the factor's
  precision; the upper triangle is never touched;
* slvtst: ||x - xc||_inf / (n eps ||A|| ||xc||_inf) <= 100 against the
  oracle's potf2 + potrs solution (atlas-tlp.base:574-630);
* info: the order of the first leading minor that is not positive definite,
  equal to the oracle potf2's, per system; other systems of the same batch
  are unaffected;
* mixed precision: the normwise backward error
  ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf) of the refined solution
  is at FP64 level; converged systems leave A and B bit-identical; the
  systems that fall back carry iter < 0 and the FP64 factor in A.
Systems follow the batched workload's generator, A = B B^T / n + I with B
iid normal(0, 1), at sizes the oracle finishes in seconds.
"""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

# unit roundoffs 2^-53 and 2^-24: half of ATL_epsilon (2^-52 in double), so the
# residual bounds below are twice as strict as the testers'
EPS64 = np.finfo(np.float64).eps / 2
EPS32 = np.finfo(np.float32).eps / 2


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def spd(n, rng):
    b = rng.standard_normal((n, n))
    return b @ b.T / max(n, 1) + np.eye(n)


class Batch:
    """Systems packed into one device buffer (column-major, per-system lda)."""

    def __init__(self, mats, lds, rhs=None, ldbs=None, dtype=np.float64):
        import torch
        self.n = [m.shape[0] for m in mats]
        self.lda = list(lds)
        self.dtype = dtype
        tdt = torch.float64 if dtype == np.float64 else torch.float32
        offs, tot = [], 0
        for n, ld in zip(self.n, self.lda):
            offs.append(tot)
            tot += max(ld * n, 1)
            tot = (tot + 31) // 32 * 32
        host = np.full(max(tot, 1), 7.25, dtype=dtype)  # fill marks untouched storage
        for m, ld, o in zip(mats, self.lda, offs):
            n = m.shape[0]
            v = host[o:o + ld * n].reshape(n, ld).T  # column-major view, ld rows
            v[:n, :] = m
        self.host0 = host.copy()
        self.offs = offs
        self.buf = torch.from_numpy(host).to("cuda")
        es = self.buf.element_size()
        self.ptrs = [self.buf.data_ptr() + o * es for o in offs]
        if rhs is not None:
            self.ldb = list(ldbs)
            boffs, btot = [], 0
            for n, ld, r in zip(self.n, self.ldb, rhs):
                boffs.append(btot)
                btot += max(ld * r.shape[1], 1)
                btot = (btot + 31) // 32 * 32
            bh = np.full(max(btot, 1), -3.5, dtype=dtype)
            for r, ld, o in zip(rhs, self.ldb, boffs):
                n, k = r.shape
                bh[o:o + ld * k].reshape(k, ld).T[:n, :] = r
            self.bhost0 = bh.copy()
            self.boffs = boffs
            self.bbuf = torch.from_numpy(bh).to("cuda")
            self.bptrs = [self.bbuf.data_ptr() + o * es for o in boffs]

    def mat(self, k, host=None):
        h = self.buf.cpu().numpy() if host is None else host
        n, ld, o = self.n[k], self.lda[k], self.offs[k]
        return np.array(h[o:o + ld * n].reshape(n, ld).T[:n, :], dtype=np.float64, order="F")

    def rhs(self, k, nrhs, host=None):
        h = self.bbuf.cpu().numpy() if host is None else host
        n, ld, o = self.n[k], self.ldb[k], self.boffs[k]
        return np.array(h[o:o + ld * nrhs].reshape(nrhs, ld).T[:n, :], dtype=np.float64,
                        order="F")


def sizes_mixed(rng, count, hi):
    edge = [1, 2, 3, 15, 16, 17, 31, 32, 33, 47, 48, 63, 64, 65, 96, 97, 127, 128, 129]
    edge = [e for e in edge if e <= hi]
    return edge + list(rng.integers(1, hi + 1, size=count - len(edge)))


def llt_ratio(a0, L, eps):
    n = a0.shape[0]
    if n == 0:
        return 0.0
    Lt = np.tril(L)
    d = np.abs(a0 - Lt @ Lt.T).sum(axis=0).max()
    return d / (np.abs(a0).sum(axis=0).max() * eps * n)


def oracle_solve(a0, b0):
    n = a0.shape[0]
    L = np.asfortranarray(a0.copy())
    assert O.potf2(O.Lower, L) == 0
    x = np.asfortranarray(b0.copy())
    O.potrs(O.Lower, L, x, n=n, nrhs=b0.shape[1])
    return x


def check_upper_untouched(batch, k, host):
    n, ld, o = batch.n[k], batch.lda[k], batch.offs[k]
    now = host[o:o + ld * n].reshape(n, ld).T
    was = batch.host0[o:o + ld * n].reshape(n, ld).T
    iu = np.triu_indices(n, 1)
    assert np.array_equal(now[:n][iu], was[:n][iu]), "upper triangle of system %d changed" % k
    if ld > n:
        assert np.array_equal(now[n:], was[n:]), "padding rows of system %d changed" % k


@pytest.mark.parametrize("nrhs", [4, 1, 7])
def test_dposv_batched_uneven(G, nrhs):
    import torch
    rng = np.random.default_rng(1000 + nrhs)
    ns = sizes_mixed(rng, 120, 200) + [0, 255, 256, 257, 384, 512]
    mats = [spd(n, rng) for n in ns]
    lds = [n + (3 if k % 3 == 0 else 0) for k, n in enumerate(ns)]
    lds = [max(l, 1) for l in lds]
    rhs = [rng.standard_normal((n, nrhs)) for n in ns]
    ldbs = [max(n + (k % 2), 1) for k, n in enumerate(ns)]
    b = Batch(mats, lds, rhs, ldbs)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    G.dposv_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, nrhs, info)
    torch.cuda.synchronize()
    inf = info.cpu().numpy()
    host, bhost = b.buf.cpu().numpy(), b.bbuf.cpu().numpy()
    worst_f = worst_s = 0.0
    for k, n in enumerate(ns):
        assert inf[k] == 0, (k, n, inf[k])
        check_upper_untouched(b, k, host)
        if n == 0:
            continue
        L = b.mat(k, host)
        worst_f = max(worst_f, llt_ratio(mats[k], L, EPS64))
        xc = oracle_solve(mats[k], rhs[k])
        x = b.rhs(k, nrhs, bhost)
        worst_s = max(worst_s, O.slv_resid(n, nrhs, O.genrm1(np.asfortranarray(mats[k])), xc,
                                           np.asfortranarray(x)))
    print("nrhs=%d systems=%d worst llt ratio %.3g, worst slv ratio %.3g"
          % (nrhs, len(ns), worst_f, worst_s))
    assert worst_f <= 16.0
    assert worst_s <= 100.0


def test_dpotrf_batched_info(G):
    """Non-SPD systems report the oracle's info; their neighbours still factor."""
    import torch
    rng = np.random.default_rng(7)
    ns, mats, want = [], [], []
    for k in range(60):
        n = int(rng.integers(2, 150))
        a = spd(n, rng)
        if k % 3 == 1:
            p = int(rng.integers(0, n))
            a[p, p] = -1.0 - a[p, p]
        ns.append(n)
        mats.append(a)
        c = np.asfortranarray(a.copy())
        want.append(O.potf2(O.Lower, c))
    b = Batch(mats, ns)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    G.dpotrf_batched(b.n, b.ptrs, b.lda, info)
    torch.cuda.synchronize()
    got = info.cpu().numpy()
    host = b.buf.cpu().numpy()
    assert list(got) == want
    for k in range(len(ns)):
        check_upper_untouched(b, k, host)
        if want[k] == 0:
            assert llt_ratio(mats[k], b.mat(k, host), EPS64) <= 16.0


def test_dposv_batched_info_keeps_rhs(G):
    """posv with non-SPD systems mixed in, on both sides of the order (512)
    up to which the forward solve runs inside the factorization: info equals
    the oracle potf2's, a failing system's right-hand sides (and every
    system's padding rows of B) are left as they were, the others solve."""
    import torch
    rng = np.random.default_rng(23)
    nrhs = 3
    ns = [5, 40, 129, 300, 511, 512, 513, 600, 64, 200, 520, 17]
    mats, want, rhs = [], [], []
    for k, n in enumerate(ns):
        a = spd(n, rng)
        if k % 3 == 1:  # a non-positive pivot part way down
            p = int(rng.integers(n // 2, n))
            a[p, p] = -1.0 - a[p, p]
        mats.append(a)
        want.append(O.potf2(O.Lower, np.asfortranarray(a.copy())))
        rhs.append(rng.standard_normal((n, nrhs)))
    ldbs = [n + 2 for n in ns]
    b = Batch(mats, ns, rhs, ldbs)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    G.dposv_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, nrhs, info)
    torch.cuda.synchronize()
    got = info.cpu().numpy()
    host, bhost = b.buf.cpu().numpy(), b.bbuf.cpu().numpy()
    assert list(got) == want
    assert any(w > 0 for w in want) and any(w == 0 for w in want)
    for k, n in enumerate(ns):
        check_upper_untouched(b, k, host)
        o, ld = b.boffs[k], b.ldb[k]
        now = bhost[o:o + ld * nrhs].reshape(nrhs, ld).T
        was = b.bhost0[o:o + ld * nrhs].reshape(nrhs, ld).T
        assert np.array_equal(now[n:], was[n:]), "padding rows of B, system %d" % k
        if want[k] > 0:
            assert np.array_equal(now, was), "B of failing system %d changed" % k
            continue
        xc = oracle_solve(mats[k], rhs[k])
        r = O.slv_resid(n, nrhs, O.genrm1(np.asfortranarray(mats[k])), xc,
                        np.asfortranarray(b.rhs(k, nrhs, bhost)))
        assert r <= 100.0, (k, n, r)


def test_dpotrs_batched_with_given_factor(G):
    """Solve-only entry with factors from the oracle."""
    import torch
    rng = np.random.default_rng(11)
    ns = [int(x) for x in rng.integers(1, 300, size=40)]
    mats, facs, rhs = [], [], []
    for n in ns:
        a = spd(n, rng)
        L = np.asfortranarray(a.copy())
        assert O.potf2(O.Lower, L) == 0
        mats.append(a)
        facs.append(L)
        rhs.append(rng.standard_normal((n, 4)))
    b = Batch(facs, ns, rhs, ns)
    G.dpotrs_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, 4)
    torch.cuda.synchronize()
    bhost = b.bbuf.cpu().numpy()
    assert np.array_equal(b.buf.cpu().numpy(), b.host0)  # factor untouched
    for k, n in enumerate(ns):
        xc = oracle_solve(mats[k], rhs[k])
        r = O.slv_resid(n, 4, O.genrm1(np.asfortranarray(mats[k])), xc,
                        np.asfortranarray(b.rhs(k, 4, bhost)))
        assert r <= 100.0, (k, n, r)


def test_many_small_systems_queue(G):
    """Thousands of tiny systems: every queue entry is taken exactly once."""
    import torch
    rng = np.random.default_rng(3)
    ns = [int(x) for x in rng.integers(1, 9, size=3000)]
    mats = [spd(n, rng) for n in ns]
    rhs = [rng.standard_normal((n, 2)) for n in ns]
    b = Batch(mats, ns, rhs, ns)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    G.dposv_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, 2, info)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    bhost = b.bbuf.cpu().numpy()
    for k in range(0, len(ns), 7):
        x = b.rhs(k, 2, bhost)
        xc = np.linalg.solve(mats[k], rhs[k])
        assert np.abs(x - xc).max() <= 1e-12 * max(1.0, np.abs(xc).max())


@pytest.mark.parametrize("t16_max", [0, 100000])
def test_tile_choice_overrides(G, t16_max):
    """Both block-column widths give tester-accurate factors on the same batch."""
    import torch
    rng = np.random.default_rng(5)
    ns = sizes_mixed(rng, 40, 300)
    mats = [spd(n, rng) for n in ns]
    old = G.batched_tune(0)
    G.set_batched_tune(0, t16_max, old[1])
    try:
        b = Batch(mats, ns)
        info = torch.zeros(len(ns), dtype=torch.int32, device="cuda")
        G.dpotrf_batched(b.n, b.ptrs, b.lda, info)
        torch.cuda.synchronize()
    finally:
        G.set_batched_tune(0, *old)
    host = b.buf.cpu().numpy()
    assert (info.cpu().numpy() == 0).all()
    for k, n in enumerate(ns):
        assert llt_ratio(mats[k], b.mat(k, host), EPS64) <= 16.0


def test_sposv_batched_fp32(G):
    import torch
    rng = np.random.default_rng(17)
    ns = sizes_mixed(rng, 60, 256)
    mats = [spd(n, rng) for n in ns]
    rhs = [rng.standard_normal((n, 4)) for n in ns]
    b = Batch([m.astype(np.float32) for m in mats], ns, [r.astype(np.float32) for r in rhs], ns,
              dtype=np.float32)
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    G.sposv_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, 4, info)
    torch.cuda.synchronize()
    assert (info.cpu().numpy() == 0).all()
    host, bhost = b.buf.cpu().numpy(), b.bbuf.cpu().numpy()
    worst_f = worst_s = 0.0
    for k, n in enumerate(ns):
        a32 = mats[k].astype(np.float32).astype(np.float64)
        worst_f = max(worst_f, llt_ratio(a32, b.mat(k, host), EPS32))
        x = b.rhs(k, 4, bhost)
        r32 = rhs[k].astype(np.float32).astype(np.float64)
        res = np.abs(r32 - a32 @ x).max(axis=0)
        den = np.abs(a32).sum(axis=1).max() * np.abs(x).max(axis=0) * n * EPS32
        worst_s = max(worst_s, (res / den).max())
    print("fp32 worst llt ratio %.3g, worst backward ratio %.3g" % (worst_f, worst_s))
    assert worst_f <= 16.0
    assert worst_s <= 100.0


def backward_error(a, x, b):
    r = np.abs(b - a @ x).max(axis=0)
    return (r / (np.abs(a).sum(axis=1).max() * np.abs(x).max(axis=0) + np.abs(b).max(axis=0))).max()


def ill_spd(n, cond, rng):
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    s = cond ** (-np.arange(n) / max(n - 1, 1))
    return (q * s) @ q.T


def test_dsposv_batched_refinement_and_fallback(G):
    import torch
    rng = np.random.default_rng(23)
    ns = sizes_mixed(rng, 50, 200)
    mats = [spd(n, rng) for n in ns]
    ill = []
    for k in range(6):  # FP32 cannot factor or refine these: FP64 fallback
        n = int(rng.integers(20, 120))
        ns.append(n)
        mats.append(ill_spd(n, 1e11, rng))
        ill.append(len(ns) - 1)
    rhs = [rng.standard_normal((n, 4)) for n in ns]
    b = Batch(mats, ns, rhs, ns)
    X = Batch(mats, ns, [np.zeros((n, 4)) for n in ns], [n + 1 for n in ns])
    info = torch.full((len(ns),), -7, dtype=torch.int32, device="cuda")
    it = torch.full((len(ns),), 99, dtype=torch.int32, device="cuda")
    G.dsposv_batched(b.n, b.ptrs, b.lda, b.bptrs, b.ldb, X.bptrs, X.ldb, 4, info, it)
    torch.cuda.synchronize()
    inf, its = info.cpu().numpy(), it.cpu().numpy()
    host, bhost, xhost = b.buf.cpu().numpy(), b.bbuf.cpu().numpy(), X.bbuf.cpu().numpy()
    assert np.array_equal(bhost, b.bhost0)  # B is input only
    worst = 0.0
    for k, n in enumerate(ns):
        assert inf[k] == 0, (k, n, inf[k])
        x = X.rhs(k, 4, xhost)
        be = backward_error(mats[k], x, rhs[k])
        worst = max(worst, be)
        assert be <= 1e-15 * max(n, 1) * 4, (k, n, be, its[k])
        if k in ill:
            assert its[k] < 0, (k, its[k])
            assert llt_ratio(mats[k], b.mat(k, host), EPS64) <= 16.0
        elif its[k] >= 0:
            assert its[k] <= 30
            a_now = b.mat(k, host)
            assert np.array_equal(a_now, mats[k]), "converged system %d: A changed" % k
        else:
            # LAPACK's test ||r|| < ||x|| ||A|| eps sqrt(n) sits at FP64
            # roundoff for the smallest orders, which may then run to the
            # iteration cap and take the FP64 path (A holds that factor)
            assert n < 8 and its[k] == -31, (k, n, its[k])
            assert llt_ratio(mats[k], b.mat(k, host), EPS64) <= 16.0
    print("dsposv: iterations %s, fallbacks %d, worst backward error %.3g"
          % (np.bincount(its[its >= 0]).tolist(), int((its < 0).sum()), worst))
