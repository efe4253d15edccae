"""GPU parity of the LAPACK entry points (clapack_dgetrf/dgetrs/dgesv,
clapack_dpotrf/dpotrs/dposv, ATL_dlaswp and the atlg_* device forms)
against the CPU oracle with ATLAS's tester formulas:

* lutst: ||P A - L U||_1 / (||A||_1 eps min(M,N)) <= 100, A = ATL_gegen with
  seed M*N+lda (AtlasBase/Clint/atlas-tlp.base:1992-2058); pivots identical
  to the oracle's right-looking LU (first-index argmax) where pivots are
  well separated;
* llttst: ||A - L L^T|| / (||A|| eps N) <= 16 (AtlasTest's -T 16), lltgen
  inputs with seed N*1029+lda (atlas-tlp.base:1320-1514), unused triangle
  untouched;
* slvtst: max ||x - xc||_inf / (N eps ||A|| ||xc||_inf) <= 100
  (atlas-tlp.base:574-630);
* info semantics: first zero pivot (continues) / first non-positive minor.
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


def lu_case(G, m, n, gap=0, check_piv=True):
    lda = max(m, 1) + gap
    a0 = O.gegen(m, n, m * n + lda, lda=lda)
    lu = a0.copy(order="F")
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int32)
    info = G.clapack_dgetrf(G.ColMajor, m, n, lu, lda, ipiv)
    mn = min(m, n)
    assert info == 0
    assert O.gechkgap(lu, m) == 0
    r = O.lib().orc_lu_resid(m, n, O._p(a0), lda, O._p(lu), lda, O._pi(ipiv))
    assert r <= 100.0, (m, n, r)
    if check_piv:
        ref = a0.copy(order="F")
        rp = np.zeros(max(mn, 1), dtype=np.int32)
        O.lib().orc_getf2(m, n, O._p(ref), lda, O._pi(rp))
        assert np.array_equal(ipiv[:mn], rp[:mn]), (m, n)
    return r


def test_getrf_lutst_sweep(G):
    for m in (1, 28, 55, 109, 163, 190):
        for n in (1, 28, 82, 136, 190):
            lu_case(G, m, n, gap=1)


def test_getrf_panel_paths(G):
    # 1000 rows: single-workgroup panels; 1500/2100 rows: multi-workgroup
    # column steps; wide and tall rectangles
    for (m, n) in ((1000, 1000), (1500, 700), (2100, 2100), (600, 1700), (3000, 64)):
        lu_case(G, m, n)


def test_getrf_large_ragged(G):
    # orders >= 2048 in block-column steps of 256 with ragged ends, both
    # rectangles, and a zero column past the first block (info = 1001)
    for (m, n) in ((2600, 2300), (2300, 2900)):
        lu_case(G, m, n, gap=3)
    n = 2300
    a0 = O.gegen(n, n, 777)
    a0[:, 1000] = 0.0
    lu = a0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv) == 1001
    assert np.all(np.isfinite(lu))
    assert O.lu_resid(n, n, a0, lu, ipiv) <= 100


def test_getrf_repeatable(G):
    # the look-ahead driver runs panels and updates on two streams over
    # disjoint columns, and no reduction depends on timing: repeated
    # factorizations of one matrix agree bit for bit
    n = 2600
    a0 = O.gegen(n, n, 4242)
    outs = []
    for _ in range(3):
        lu = a0.copy(order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        assert G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv) == 0
        outs.append((lu, ipiv))
    for lu, ipiv in outs[1:]:
        assert np.array_equal(lu, outs[0][0])
        assert np.array_equal(ipiv, outs[0][1])


def test_getrf_concurrent_host_threads(G):
    # two host threads factor different matrices on their own streams at
    # the same time (the look-ahead drivers share auxiliary streams and
    # events); each result equals the one computed alone
    import threading
    import torch
    n = 2304
    mats = [torch.from_numpy(O.gegen(n, n, 900 + t).T.copy()).cuda() for t in range(2)]
    alone = []
    for a0 in mats:
        w = a0.clone()
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        G.dgetrf(n, n, w, n, ipiv, info)
        torch.cuda.synchronize()
        alone.append((w, ipiv))
    for _ in range(3):
        outs = [None, None]
        errs = []

        def work(t):
            try:
                st = torch.cuda.Stream()
                with torch.cuda.stream(st):
                    w = mats[t].clone()
                    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
                    info = torch.zeros(1, dtype=torch.int32, device="cuda")
                    G.dgetrf(n, n, w, n, ipiv, info, stream=st)
                st.synchronize()
                outs[t] = (w, ipiv, int(info.item()))
            except Exception as e:  # reported below, on the main thread
                errs.append(e)

        th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
        for x in th:
            x.start()
        for x in th:
            x.join()
        assert not errs, errs
        for t in range(2):
            assert outs[t][2] == 0
            assert torch.equal(outs[t][0], alone[t][0])
            assert torch.equal(outs[t][1], alone[t][1])


@pytest.mark.parametrize("n", [24576, 32768])
def test_getrf_wide_lookahead_blocks(G, n):
    # orders whose look-ahead blocks are 768 and 1024 columns wide, checked
    # by size-independent properties on the GPU: lutst's scaled residual
    # ||P A - L U||_1 / (||A||_1 n eps) <= 100 (L U and P A formed with
    # torch in FP64), pivots in range, and |l_ij| <= 1 (partial pivoting
    # takes the largest remaining entry of each column)
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(n)
    t = torch.empty((n, n), dtype=torch.float64, device="cuda").uniform_(-1, 1, generator=g)
    w = t.clone()  # row-major t is the column-major matrix t^T
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.dgetrf(n, n, w, n, ipiv, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    p = ipiv.cpu().numpy()
    assert np.all(p >= np.arange(n)) and np.all(p < n)
    f = w.t()
    low = torch.tril(f, -1)
    assert float(low.abs().max()) <= 1.0
    low.diagonal().fill_(1.0)
    up = torch.triu(f)
    del w, f
    lu = torch.matmul(low, up)
    del low, up
    perm = np.arange(n)
    for i in range(n):
        perm[i], perm[p[i]] = perm[p[i]], perm[i]
    a = t.t()
    lu.sub_(a[torch.from_numpy(perm).cuda()])
    r = float(lu.abs().sum(dim=0).max()) / (float(a.abs().sum(dim=0).max()) * n * 2.0 ** -53)
    print("getrf n=%d: lutst ratio %.3g" % (n, r))
    assert r <= 100.0


def test_getrf_known_pivots(G):
    # A = P^T L U with |L_ij| < 0.45/n and diag(U) in [1, 2]: at every step the
    # unit-diagonal row beats every competitor by a factor > n (and L, U are
    # well conditioned, so rounding in forming A cannot change that)
    rng = np.random.default_rng(3)
    for n in (700, 2100):
        L = np.tril(rng.uniform(-0.45, 0.45, (n, n)), -1) / n + np.eye(n)
        U = np.triu(rng.uniform(-1, 1, (n, n)), 1) / n + np.diag(rng.uniform(1, 2, n))
        perm = rng.permutation(n)
        a = np.asfortranarray((L @ U)[np.argsort(perm)])
        lu = a.copy(order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        assert G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv) == 0
        rows = np.arange(n)
        for i, q in enumerate(ipiv):
            rows[[i, q]] = rows[[q, i]]
        assert np.array_equal(np.argsort(perm)[rows], np.arange(n))
        assert O.lu_resid(n, n, a, lu, ipiv) <= 100


def test_getrf_zero_column_info(G):
    n = 300
    a0 = O.gegen(n, n, 4242)
    a0[:, 37] = 0.0
    lu = a0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv) == 38
    assert np.all(np.isfinite(lu))
    assert O.lu_resid(n, n, a0, lu, ipiv) <= 100
    # zero column in a tall multi-workgroup panel
    m, n = 2000, 40
    a0 = O.gegen(m, n, 99)
    a0[:, 5] = 0.0
    lu = a0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert G.clapack_dgetrf(G.ColMajor, m, n, lu, m, ipiv) == 6
    assert O.lu_resid(m, n, a0, lu, ipiv) <= 100


def test_getrf_rowmajor(G):
    # row-major M x N is column-major N x M of the same memory (ATL_getrfR)
    m, n = 150, 220
    a0 = O.gegen(n, m, m * n + n)     # memory: column-major n x m
    lu = a0.copy(order="F")
    ipiv = np.zeros(min(m, n), dtype=np.int32)
    assert G.clapack_dgetrf(G.RowMajor, m, n, lu, n, ipiv) == 0
    ref = a0.copy(order="F")
    rp, _ = O.getf2(ref)
    assert np.array_equal(ipiv, rp)
    assert O.lu_resid(n, m, a0, lu, ipiv) <= 100


def test_laswp_against_oracle(G):
    rng = np.random.default_rng(5)
    m, ncol = 900, 37
    a0 = np.asfortranarray(rng.standard_normal((m, ncol)))
    for k1, k2, inc in ((0, 700, 1), (0, 700, -1), (100, 400, 1), (0, 900, -1)):
        ipiv = rng.integers(0, m, size=m).astype(np.int32)  # arbitrary, not only >= i
        got = a0.copy(order="F")
        G.ATL_dlaswp(ncol, got, m, k1, k2, ipiv, inc)
        ref = a0.copy(order="F")
        O.laswp(ref, k1, k2, ipiv, inc)
        assert np.array_equal(got, ref), (k1, k2, inc)


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_potrf_llttst(G, uplo):
    for n in (1, 28, 82, 136, 190, 300, 1000):
        lda = n + 1
        s = O.lltgen(uplo, n, n * 1029 + lda, lda=lda)
        # sentinel in the unused triangle: must survive untouched
        mask = np.triu(np.ones((n, n), bool), 1) if uplo == O.Lower else np.tril(np.ones((n, n), bool), -1)
        s[:n][mask] = -973200000.0
        f = s.copy(order="F")
        assert G.clapack_dpotrf(G.ColMajor, uplo, n, f, lda) == 0
        assert np.all(f[:n][mask] == -973200000.0)
        assert O.gechkgap(f, n) == 0
        r = O.lib().orc_llt_resid(uplo, n, O._p(s), lda, O._p(f), lda)
        assert r <= 16.0, (n, r)


def test_potrf_info_indefinite(G):
    for n, k in ((300, 157), (1000, 700), (200, 0)):
        s = O.lltgen(O.Lower, n, 12345)
        s[k, k] = -1.0e6
        f = s.copy(order="F")
        assert G.clapack_dpotrf(G.ColMajor, G.Lower, n, f, n) == k + 1


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_potrf_large_ragged(G, uplo):
    # orders >= 2048 (block-column steps of 256, ragged last block): llttst,
    # the unused triangle untouched, and info from a minor in a later block
    n = 2300
    lda = n + 3
    s = O.lltgen(uplo, n, n * 1029 + lda, lda=lda)
    mask = np.triu(np.ones((n, n), bool), 1) if uplo == O.Lower else np.tril(np.ones((n, n), bool), -1)
    s[:n][mask] = -973200000.0
    f = s.copy(order="F")
    assert G.clapack_dpotrf(G.ColMajor, uplo, n, f, lda) == 0
    assert np.all(f[:n][mask] == -973200000.0)
    assert O.gechkgap(f, n) == 0
    r = O.lib().orc_llt_resid(uplo, n, O._p(s), lda, O._p(f), lda)
    assert r <= 16.0, (n, r)
    s2 = O.lltgen(uplo, n, 777)
    s2[1700, 1700] = -1.0e6
    f2 = s2.copy(order="F")
    assert G.clapack_dpotrf(G.ColMajor, uplo, n, f2, n) == 1701


def slv_inputs(n, nrhs, lda, ldb):
    a = O.gegen(n, n, n * n + lda, lda=lda)
    x = O.gegen(n, nrhs, n * nrhs + ldb, lda=ldb)
    b = np.zeros((ldb, nrhs), order="F")
    O.lib().orc_refgemm(O.NoTrans, O.NoTrans, n, nrhs, n, 1.0, O._p(a), lda, O._p(x), ldb, 0.0,
                        O._p(b), ldb)
    return a, x, b


def test_gesv_getrs_slvtst(G):
    for n, nrhs in ((167, 83), (500, 1), (1300, 4)):
        lda, ldb = n + 2, n + 3
        a, x, b = slv_inputs(n, nrhs, lda, ldb)
        nrm = O.genrm1(a, n, n)
        lu = a.copy(order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        bb = b.copy(order="F")
        assert G.clapack_dgesv(G.ColMajor, n, nrhs, lu, lda, ipiv, bb, ldb) == 0
        r = O.lib().orc_slv_resid(n, nrhs, nrm, O._p(x), ldb, O._p(bb), ldb)
        assert r <= 100, (n, nrhs, r)
        # transposed solve with the same factors: A^T x = A^T xc
        bt = np.zeros((ldb, nrhs), order="F")
        O.lib().orc_refgemm(O.Trans, O.NoTrans, n, nrhs, n, 1.0, O._p(a), lda, O._p(x), ldb, 0.0,
                            O._p(bt), ldb)
        assert G.clapack_dgetrs(G.ColMajor, G.Trans, n, nrhs, lu, lda, ipiv, bt, ldb) == 0
        r = O.lib().orc_slv_resid(n, nrhs, nrm, O._p(x), ldb, O._p(bt), ldb)
        assert r <= 100, (n, nrhs, r)


def test_gesv_rowmajor(G):
    # row-major A (n x n) is column-major A^T; B holds NRHS rows of length n
    n, nrhs = 240, 3
    a = O.gegen(n, n, 31)          # memory = row-major A with A = a.T
    x = O.gegen(n, nrhs, 32)       # column j = right-hand side j
    b = np.asfortranarray(a.T @ x)
    lu = a.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert G.clapack_dgesv(G.RowMajor, n, nrhs, lu, n, ipiv, b, n) == 0
    assert O.slv_resid(n, nrhs, O.genrm1(np.asfortranarray(a.T)), x, b) <= 100


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_posv_potrs_slvtst(G, uplo):
    for n, nrhs in ((167, 83), (700, 4)):
        s = O.lltgen(uplo, n, n * 1029 + n)
        full = np.tril(s) + np.tril(s, -1).T if uplo == O.Lower else np.triu(s) + np.triu(s, 1).T
        x = O.gegen(n, nrhs, n * nrhs + n)
        b = np.asfortranarray(full @ x)
        f = s.copy(order="F")
        bb = b.copy(order="F")
        assert G.clapack_dposv(G.ColMajor, uplo, n, nrhs, f, n, bb, n) == 0
        nrm = np.abs(full).sum(axis=0).max()
        assert O.slv_resid(n, nrhs, nrm, x, bb) <= 100
        for order in (G.ColMajor, G.RowMajor):
            bb = b.copy(order="F")
            # row-major uplo is the column-major opposite of the same memory
            u2 = uplo if order == G.ColMajor else (G.Upper if uplo == G.Lower else G.Lower)
            assert G.clapack_dpotrs(order, u2, n, nrhs, f, n, bb, n) == 0
            assert O.slv_resid(n, nrhs, nrm, x, bb) <= 100


def test_device_api_getrf_potrf(G):
    import torch
    n = 1500
    a0 = O.gegen(n, n, n * n + n)
    da = G.from_fortran(a0)
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.dgetrf(n, n, da, n, ipiv, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert O.lu_resid(n, n, a0, G.to_fortran(da), ipiv.cpu().numpy()) <= 100
    s = O.lltgen(O.Lower, n, n * 1029 + n)
    ds = G.from_fortran(s)
    G.dpotrf(G.Lower, n, ds, n, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    assert O.llt_resid(O.Lower, n, s, G.to_fortran(ds)) <= 16
