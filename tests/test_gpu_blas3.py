"""GPU parity of the level-3 BLAS entry points (cblas_dgemm, cblas_dsyrk,
cblas_dtrsm and their atlg_* device forms) against the CPU oracle, judged
with ATLAS's own tester formulas:

* l3blastst GEMM/TRSM residual < 50 (AtlasBase/Antoine/atlbinblas.base:
  11791-11935, 12761-12900), inputs from ATL_gegen with l3blastst's seeds;
* gemmtst element-wise bound |D-C| <= 2(|alpha| 2K eps + |beta| eps) + eps
  (AtlasBase/Clint/atlas.base:21968);
* lda-gap sentinels untouched (gechkgap), beta == 0 overwrites NaN in C.

Sizes follow ATLAS's test data: the netlib c_dblat3.dat set
(N in {0,1,2,3,5,9}, alpha in {0,1,0.7}, beta in {0,1,1.3}) and the AtlasTest
binrun sweep 1..195 step 27 with alpha, beta in {0, 1, -1.1}.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = json.load(open(os.path.join(HERE, "golden", "atlas_test_params.json")))
SWEEP = [1, 28, 55, 82, 109, 136, 163, 190]


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    atlas_gpu.lib()
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def gemm_case(G, ta, tb, m, n, k, alpha, beta, gap=0, order=None, nan_c=False, even=False):
    mA, nA = (m, k) if ta == O.NoTrans else (k, m)
    mB, nB = (k, n) if tb == O.NoTrans else (n, k)
    lda, ldb, ldc = max(mA, 1) + gap, max(mB, 1) + gap, max(m, 1) + gap
    if even:  # A and B 16-byte aligned column by column (C keeps its gap)
        lda, ldb = lda + (lda & 1), ldb + (ldb & 1)
    a = O.gegen(mA, nA, mA * nA + 513 * 7 + 90, lda=lda)
    b = O.gegen(mB, nB, mB * nB + 127 * 50 + 77, lda=ldb)
    c0 = O.gegen(m, n, m * n + 101 * 2 + 53, lda=ldc)
    if nan_c:
        c0[:m, :n] = np.nan
    ct = c0.copy(order="F")
    if order == G.RowMajor:
        # row-major C = op(A) op(B) is column-major C^T = op(B)^T op(A)^T
        G.cblas_dgemm(G.RowMajor, tb, ta, n, m, k, alpha, b, ldb, a, lda, beta, ct, ldc)
    else:
        G.cblas_dgemm(G.ColMajor, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, ct, ldc)
    cr = c0.copy(order="F")
    O.lib().orc_refgemm(ta, tb, m, n, k, alpha, O._p(a), lda, O._p(b), ldb, beta, O._p(cr), ldc)
    if m and n:
        assert O.gechkgap(ct, m) == 0
    if nan_c:
        assert not np.any(np.isnan(ct[:m, :n]))
        c0 = np.zeros_like(c0, order="F")
    r = O.gemm_resid(ta, tb, m, n, k, alpha, a, b, beta, c0, ct, cr) if m and n else 0.0
    nbad, mx = O.gemm_elem_check(m, n, k, alpha, beta, ct, cr) if m and n else (0, 0.0)
    return r, nbad, mx


@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("tb", [O.NoTrans, O.Trans])
def test_gemm_cblat3_set(G, ta, tb):
    p = PARAMS["cblat3"]
    for n in p["N"]:
        for alpha in p["alpha"]:
            for beta in p["beta"]:
                for order in (G.ColMajor, G.RowMajor):
                    r, nbad, _ = gemm_case(G, ta, tb, n, n, n, alpha, beta, order=order)
                    assert r < 16.0 and nbad == 0, (n, alpha, beta, order, r, nbad)


@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("tb", [O.NoTrans, O.Trans])
def test_gemm_binrun_sweep(G, ta, tb):
    rng = np.random.default_rng(1)
    combos = [(m, n, k) for m in SWEEP for n in SWEEP for k in SWEEP]
    pick = rng.choice(len(combos), 24, replace=False)
    for idx in pick:
        m, n, k = combos[idx]
        for alpha, beta in ((1.0, 1.0), (-1.1, 0.0), (0.0, -1.1), (1.0, -1.1)):
            r, nbad, mx = gemm_case(G, ta, tb, m, n, k, alpha, beta, gap=3)
            assert r < 50.0 and nbad == 0, (m, n, k, alpha, beta, r, nbad, mx)


def test_gemm_beta0_overwrites_nan_and_alpha0(G):
    for (m, n, k) in ((190, 163, 55), (129, 257, 17)):
        r, nbad, _ = gemm_case(G, O.NoTrans, O.NoTrans, m, n, k, 1.0, 0.0, nan_c=True)
        assert r < 50 and nbad == 0
    r, nbad, _ = gemm_case(G, O.NoTrans, O.Trans, 100, 90, 80, 0.0, 0.0, nan_c=True)
    assert nbad == 0


def test_gemm_nn_register_path_edges(G):
    # NoTrans x NoTrans with even leading dimensions and K a multiple of 32
    # takes the kernel that loads A straight into registers: partial row
    # and column tiles (odd M included), C with an odd leading dimension
    # (scalar epilogue), beta = 0 over NaN, alpha = -1.1, and ragged grids
    for (m, n, k, gap, alpha, beta, nan_c) in (
            (129, 257, 64, 2, 1.0, 1.0, False), (127, 3, 32, 0, -1.1, 1.3, False),
            (2113, 200, 96, 4, 1.0, 0.0, True), (3, 130, 32, 1, 0.7, 1.0, False),
            (1, 1, 32, 2, 1.0, -1.1, False), (384, 384, 512, 0, 1.0, 1.0, False),
            (255, 129, 160, 2, -1.1, 0.0, True)):
        r, nbad, mx = gemm_case(G, O.NoTrans, O.NoTrans, m, n, k, alpha, beta, gap=gap,
                                nan_c=nan_c, even=True)
        assert r < 50.0 and nbad == 0, (m, n, k, gap, alpha, beta, r, nbad, mx)


@pytest.mark.parametrize("ta,tb", [(O.NoTrans, O.Trans), (O.Trans, O.NoTrans)])
def test_gemm_nt_tn_register_path_edges(G, ta, tb):
    # NoTrans x Trans and Trans x NoTrans with even leading dimensions and K
    # a multiple of 32: the register-A kernel with B^T staged row by row
    # (pair loads clamped at an odd N), or with A^T's k runs loaded per row
    # (rows clamped at M-1); the same edge cases as the NoTrans x NoTrans form
    for (m, n, k, gap, alpha, beta, nan_c) in (
            (129, 257, 64, 2, 1.0, 1.0, False), (127, 3, 32, 0, -1.1, 1.3, False),
            (2113, 201, 96, 4, 1.0, 0.0, True), (3, 131, 32, 1, 0.7, 1.0, False),
            (1, 1, 32, 2, 1.0, -1.1, False), (384, 385, 512, 0, 1.0, 1.0, False),
            (255, 129, 160, 2, -1.1, 0.0, True)):
        r, nbad, mx = gemm_case(G, ta, tb, m, n, k, alpha, beta, gap=gap, nan_c=nan_c,
                                even=True)
        assert r < 50.0 and nbad == 0, (m, n, k, gap, alpha, beta, r, nbad, mx)


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("trans", [O.NoTrans, O.Trans])
def test_syrk_register_path(G, uplo, trans):
    # SYRK with an even lda and K a multiple of 32 runs the register-A
    # kernel over the triangle's tiles only; the diagonal tiles split 16-byte
    # row pairs between kept and untouched elements
    for (n, k, ldc_gap) in ((257, 64, 2), (129, 32, 1), (600, 96, 0), (3, 32, 1), (385, 160, 3)):
        na, ka = (n, k) if trans == O.NoTrans else (k, n)
        a = O.gegen(na + (na & 1), ka, n * k + 3)  # lda even
        c0 = O.gegen(n, n, n + 17, lda=n + ldc_gap)
        for alpha, beta in ((1.0, 1.0), (-1.1, 0.0), (0.7, 1.3)):
            ct = c0.copy(order="F")
            G.cblas_dsyrk(G.ColMajor, uplo, trans, n, k, alpha, a, a.shape[0], beta, ct,
                          n + ldc_gap)
            cr = c0.copy(order="F")
            O.refsyrk(uplo, trans, n, k, alpha, np.asfortranarray(a[:na]), beta, cr)
            tri = np.tril if uplo == O.Lower else np.triu
            other = np.triu if uplo == O.Lower else np.tril
            off = 1 if uplo == O.Lower else -1
            assert np.array_equal(other(ct[:n], off), other(c0[:n], off))
            assert O.gechkgap(ct, n) == 0
            err = np.abs(tri(ct[:n]) - tri(cr[:n])).max()
            bound = 2 * (abs(alpha) * 2 * k * 2.0 ** -52 + abs(beta) * 2.0 ** -52) + 2.0 ** -52
            assert err <= bound, (n, k, alpha, beta, err)


def test_gemm_config1_1000(G):
    # the CPU-runnable headline case: M=N=K=1000, alpha=beta=1
    r, nbad, mx = gemm_case(G, O.NoTrans, O.NoTrans, 1000, 1000, 1000, 1.0, 1.0)
    assert r < 50 and nbad == 0, (r, nbad, mx)


def test_gemm_device_api_tall_skinny(G):
    import torch
    # M=65536, N=K=256 at full size; C rows checked on a sample against the oracle
    m, n, k = 65536, 256, 256
    a = O.gegen(m, k, 11)
    b = O.gegen(k, n, 12)
    c0 = O.gegen(m, n, 13)
    da, db, dc = G.from_fortran(a), G.from_fortran(b), G.from_fortran(c0)
    G.dgemm(G.NoTrans, G.NoTrans, m, n, k, 1.0, da, m, db, k, 1.0, dc, m)
    torch.cuda.synchronize()
    ct = G.to_fortran(dc)
    rows = np.r_[0:64, 30000:30064, m - 64:m]
    asub = np.asfortranarray(a[rows])
    csub = np.asfortranarray(c0[rows])
    cr = csub.copy(order="F")
    O.refgemm(O.NoTrans, O.NoTrans, len(rows), n, k, 1.0, asub, b, 1.0, cr)
    got = np.asfortranarray(ct[rows])
    assert O.gemm_elem_check(len(rows), n, k, 1.0, 1.0, got, cr)[0] == 0
    assert O.gemm_resid(O.NoTrans, O.NoTrans, len(rows), n, k, 1.0, asub, b, 1.0, csub, got, cr) < 50


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("trans", [O.NoTrans, O.Trans])
def test_syrk(G, uplo, trans):
    for (n, k) in ((1, 1), (55, 28), (190, 136), (300, 129)):
        na, ka = (n, k) if trans == O.NoTrans else (k, n)
        a = O.gegen(na, ka, na * ka + 7)
        c0 = O.gegen(n, n, n * n + 9, lda=n + 2)
        for alpha, beta in ((1.0, 1.0), (-1.1, 0.0), (0.7, 1.3)):
            ct = c0.copy(order="F")
            G.cblas_dsyrk(G.ColMajor, uplo, trans, n, k, alpha, a, na, beta, ct, n + 2)
            cr = c0.copy(order="F")
            O.refsyrk(uplo, trans, n, k, alpha, a, beta, cr)
            tri = np.tril if uplo == O.Lower else np.triu
            other = np.triu if uplo == O.Lower else np.tril
            off = 1 if uplo == O.Lower else -1
            assert np.array_equal(other(ct[:n], off), other(c0[:n], off))
            assert O.gechkgap(ct, n) == 0
            err = np.abs(tri(ct[:n]) - tri(cr[:n])).max()
            bound = 2 * (abs(alpha) * 2 * k * 2.0 ** -52 + abs(beta) * 2.0 ** -52) + 2.0 ** -52
            assert err <= bound, (n, k, alpha, beta, err)


def test_dgemm_a_equals_b_syrk_route(G):
    # cblas_dgemm routes A==B, TA!=TB, beta==0 to SYRK + reflection
    n, k = 140, 70
    a = O.gegen(n, k, 5)
    c = np.full((n, n), np.nan, order="F")
    G.cblas_dgemm(G.ColMajor, G.NoTrans, G.Trans, n, n, k, 1.0, a, n, a, n, 0.0, c, n)
    cr = np.zeros((n, n), order="F")
    O.refgemm(O.NoTrans, O.Trans, n, n, k, 1.0, a, a, 0.0, cr)
    assert np.array_equal(c, c.T)
    assert O.gemm_elem_check(n, n, k, 1.0, 0.0, c, cr)[0] == 0


SIDES = [O.Left, O.Right]
UPLOS = [O.Lower, O.Upper]
TRANS = [O.NoTrans, O.Trans]
DIAGS = [O.NonUnit, O.Unit]


@pytest.mark.parametrize("side", SIDES)
@pytest.mark.parametrize("uplo", UPLOS)
@pytest.mark.parametrize("ta", TRANS)
@pytest.mark.parametrize("diag", DIAGS)
def test_trsm_l3blastst(G, side, uplo, ta, diag):
    for (m, n) in ((1, 1), (7, 5), (128, 64), (129, 33), (300, 257), (513, 100)):
        for alpha in (1.0, -1.1, 0.0):
            na = m if side == O.Left else n
            lda, ldb = na + 2, m + 1
            a = O.gegen(na, na, na * na + 513 * 7 + 90, lda=lda)
            O.lib().orc_trddom(uplo, na, O._p(a), lda)
            if diag == O.Unit:
                a[np.arange(na), np.arange(na)] = np.nan   # never read
            b0 = O.gegen(m, n, m * n + 127 * 50 + 77, lda=ldb)
            bt = b0.copy(order="F")
            G.cblas_dtrsm(G.ColMajor, side, uplo, ta, diag, m, n, alpha, a, lda, bt, ldb)
            br = b0.copy(order="F")
            O.lib().orc_reftrsm(side, uplo, ta, diag, m, n, alpha, O._p(a), lda, O._p(br), ldb)
            assert O.gechkgap(bt, m) == 0
            assert np.all(np.isfinite(bt[:m]))
            r = O.lib().orc_trsm_resid(side, uplo, diag, m, n, alpha, O._p(a), lda, O._p(b0),
                                       ldb, O._p(bt), ldb, O._p(br), ldb)
            assert r < 50.0, (m, n, alpha, r)


@pytest.mark.parametrize("uplo", UPLOS)
@pytest.mark.parametrize("diag", DIAGS)
def test_trsm_few_rhs_long_triangle(G, uplo, diag):
    # Left/NoTrans with 1..4 right-hand sides on triangles >= 1024 (the
    # sync-free block solve), ragged last block, lda/ldb gaps, alpha != 1
    for (m, n) in ((1024, 1), (1500, 2), (2600, 3), (1111, 4)):
        for alpha in (1.0, -1.1):
            lda, ldb = m + 2, m + 1
            a = O.gegen(m, m, m * m + 513 * 7 + 90, lda=lda)
            O.lib().orc_trddom(uplo, m, O._p(a), lda)
            if diag == O.Unit:
                a[np.arange(m), np.arange(m)] = np.nan   # never read
            b0 = O.gegen(m, n, m * n + 127 * 50 + 77, lda=ldb)
            bt = b0.copy(order="F")
            G.cblas_dtrsm(G.ColMajor, O.Left, uplo, O.NoTrans, diag, m, n, alpha, a, lda, bt, ldb)
            br = b0.copy(order="F")
            O.lib().orc_reftrsm(O.Left, uplo, O.NoTrans, diag, m, n, alpha, O._p(a), lda,
                                O._p(br), ldb)
            assert O.gechkgap(bt, m) == 0
            r = O.lib().orc_trsm_resid(O.Left, uplo, diag, m, n, alpha, O._p(a), lda, O._p(b0),
                                       ldb, O._p(bt), ldb, O._p(br), ldb)
            assert r < 50.0, (m, n, alpha, r)


@pytest.mark.parametrize("uplo", UPLOS)
@pytest.mark.parametrize("diag", DIAGS)
def test_trsm_few_rhs_long_triangle_trans(G, uplo, diag):
    # Left/Trans with 1..4 right-hand sides on triangles >= 1024: the same
    # sync-free block solve reading A's blocks transposed (each thread's
    # entries contiguous in A), blocks taken in the opposite order
    for (m, n) in ((1024, 1), (1500, 2), (2600, 3), (1111, 4)):
        for alpha in (1.0, -1.1):
            lda, ldb = m + 2, m + 1
            a = O.gegen(m, m, m * m + 513 * 3 + 91, lda=lda)
            O.lib().orc_trddom(uplo, m, O._p(a), lda)
            if diag == O.Unit:
                a[np.arange(m), np.arange(m)] = np.nan   # never read
            b0 = O.gegen(m, n, m * n + 127 * 51 + 77, lda=ldb)
            bt = b0.copy(order="F")
            G.cblas_dtrsm(G.ColMajor, O.Left, uplo, O.Trans, diag, m, n, alpha, a, lda, bt, ldb)
            br = b0.copy(order="F")
            O.lib().orc_reftrsm(O.Left, uplo, O.Trans, diag, m, n, alpha, O._p(a), lda,
                                O._p(br), ldb)
            assert O.gechkgap(bt, m) == 0
            err = np.abs(bt[:m] - br[:m]).max() / np.abs(br[:m]).max()
            assert err < 1e-13, (m, n, alpha, err)


def test_trsm_rowmajor_and_device_api(G):
    import torch
    m, n = 300, 200
    a = O.gegen(m, m, 77)
    O.trddom(a, O.Lower, m)
    b0 = O.gegen(m, n, 78)
    # device path
    da, db = G.from_fortran(a), G.from_fortran(b0)
    G.dtrsm(G.Left, G.Lower, G.NoTrans, G.Unit, m, n, 1.0, da, m, db, m)
    torch.cuda.synchronize()
    br = b0.copy(order="F")
    O.reftrsm(O.Left, O.Lower, O.NoTrans, O.Unit, m, n, 1.0, a, br)
    r = O.trsm_resid(O.Left, O.Lower, O.Unit, m, n, 1.0, a, b0, G.to_fortran(db), br)
    assert r < 50
    # row-major: B is n x m row-major (= B^T column-major), X A = B with A
    # row-major upper (= column-major lower transposed)
    bt = np.asfortranarray(b0.T.copy())          # memory of row-major m x n
    x = np.ascontiguousarray(b0)                 # row-major m x n of b0
    xr = np.asfortranarray(b0.copy())
    # left/lower/notrans in row-major on a row-major A = column-major A^T
    arow = np.asfortranarray(a.T.copy())        # memory holding row-major a
    G.cblas_dtrsm(G.RowMajor, G.Left, G.Lower, G.NoTrans, G.Unit, m, n, 1.0, arow, m, bt, n)
    got = np.asfortranarray(bt.T)
    O.reftrsm(O.Left, O.Lower, O.NoTrans, O.Unit, m, n, 1.0, a, xr)
    assert O.trsm_resid(O.Left, O.Lower, O.Unit, m, n, 1.0, a, b0, got, xr) < 50
    del x
