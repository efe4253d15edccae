"""Pins the CPU oracle (oracle/atl_oracle.c) before anything is checked
against it.  Each oracle function is tied here to something it did not
produce itself: the closed-form LCG (tests/golden/lcg_words.json), exact
small-integer arithmetic, numpy/scipy (LAPACK) results, invariants such as
P*A = L*U, and planted mistakes that must move a residual.

ATLAS ships no stored golden vectors for this path; its testers generate
inputs with the LCG and check scaled residuals, so these pins cover the
generator, the residual formulas and the textbook routines separately.
"""
import json
import os

import numpy as np
import pytest
import scipy.linalg as sla

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def rng(seed=0):
    return np.random.default_rng(seed)


def small_int_mat(r, m, n, lo=-4, hi=5):
    return np.asfortranarray(r.integers(lo, hi, size=(m, n)).astype(np.float64))


# ---------------------------------------------------------------- generator

def test_lcg_matches_closed_form_words(orc):
    with open(os.path.join(GOLD, "lcg_words.json")) as f:
        words = json.load(f)["words"]
    for seed, ws in words.items():
        got = O.rand_words(int(seed), len(ws))
        assert list(got) == ws, seed


def test_lcg_is_the_31bit_ansi_lcg(orc):
    # independent restatement: X(n+1) = 1103515245 X(n) + 12345 mod 2^31
    w = O.rand_words(12345, 1000).astype(np.int64)
    assert np.all((w >= 0) & (w < 2 ** 31))
    assert np.all((1103515245 * w[:-1] + 12345) % 2 ** 31 == w[1:])


def test_gegen_range_order_and_gap(orc):
    m, n, lda, seed = 7, 5, 10, 1234
    a = O.gegen(m, n, seed, lda=lda)
    w = O.rand_words(seed, m * n).astype(np.float64)
    expect = (0.5 - w / 2.0 ** 31).reshape(n, m).T  # column by column
    assert np.array_equal(a[:m, :], expect)
    assert np.all(a[:m] > -0.5) and np.all(a[:m] <= 0.5)
    assert np.all(a[m:, :] == -2560000000.0)
    assert O.gechkgap(a, m) == 0
    a[m + 1, 2] = 0.0
    assert O.gechkgap(a, m) == 1


def test_epsilon_closed_form(orc):
    assert O.epsilon() == 2.0 ** -52 == np.finfo(np.float64).eps


# ---------------------------------------------------------------- norms

def test_norms_against_numpy(orc):
    r = rng(1)
    a = np.asfortranarray(r.standard_normal((9, 6)))
    b = np.asfortranarray(r.standard_normal((9, 6)))
    assert O.genrm1(a) == pytest.approx(np.abs(a).sum(axis=0).max(), rel=1e-15)
    d = O.lib().orc_gediffnrm1(9, 6, O._p(a), 9, O._p(b), 9)
    assert d == pytest.approx(np.abs(a - b).sum(axis=0).max(), rel=1e-15)
    t = np.asfortranarray(r.standard_normal((6, 6)))
    for uplo, tri in ((O.Lower, np.tril), (O.Upper, np.triu)):
        nu = O.trnrm1(uplo, O.NonUnit, t, 6)
        assert nu == pytest.approx(np.abs(tri(t)).sum(axis=0).max(), rel=1e-15)
        tu = tri(t, -1 if uplo == O.Lower else 1) + np.eye(6)
        assert O.trnrm1(uplo, O.Unit, t, 6) == pytest.approx(np.abs(tu).sum(axis=0).max(), rel=1e-15)
        # lltnrm1: 1-norm of the symmetric matrix stored in one triangle
        s = tri(t)
        full = s + s.T - np.diag(np.diag(s))
        assert O.lltnrm1(uplo, np.asfortranarray(s), 6) == pytest.approx(
            np.abs(full).sum(axis=0).max(), rel=1e-15)
    nan = a.copy(order="F")
    nan[3, 2] = np.nan
    assert np.isnan(O.genrm1(nan))


def test_trddom_and_lltgen(orc):
    n = 6
    a = O.gegen(n, n, 99)
    a0 = a.copy(order="F")
    O.trddom(a, O.Lower, n)
    assert np.array_equal(np.tril(a, -1), np.tril(a0, -1) * (1.0 / n))  # multiply by 1/N
    d0 = np.diag(a0)
    assert np.array_equal(np.diag(a), np.where(d0 >= 0, d0 + 1, d0 - 1))
    l = O.lltgen(O.Lower, n, n * 1029 + n)
    assert np.all(np.triu(l, 1) == 0)
    full = np.tril(l) + np.tril(l, -1).T
    np.linalg.cholesky(full)  # positive definite by diagonal dominance


# ---------------------------------------------------------------- GEMM / SYRK

@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("tb", [O.NoTrans, O.Trans])
def test_refgemm_exact_small_ints(orc, ta, tb):
    r = rng(2)
    m, n, k = 7, 5, 6
    a = small_int_mat(r, m if ta == O.NoTrans else k, k if ta == O.NoTrans else m)
    b = small_int_mat(r, k if tb == O.NoTrans else n, n if tb == O.NoTrans else k)
    c = small_int_mat(r, m, n)
    opa = a if ta == O.NoTrans else a.T
    opb = b if tb == O.NoTrans else b.T
    for alpha, beta in ((1.0, 1.0), (2.0, -3.0), (0.0, 2.0), (-1.0, 0.0)):
        cc = c.copy(order="F")
        O.refgemm(ta, tb, m, n, k, alpha, a, b, beta, cc)
        assert np.array_equal(cc, alpha * opa @ opb + beta * c)


def test_refgemm_beta0_overwrites_nan(orc):
    r = rng(3)
    a, b = small_int_mat(r, 4, 3), small_int_mat(r, 3, 5)
    c = np.full((4, 5), np.nan, order="F")
    O.refgemm(O.NoTrans, O.NoTrans, 4, 5, 3, 1.0, a, b, 0.0, c)
    assert np.array_equal(c, a @ b)
    c = np.full((4, 5), np.nan, order="F")
    O.refgemm(O.NoTrans, O.NoTrans, 4, 5, 3, 0.0, a, b, 0.0, c)
    assert np.all(c == 0)


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("trans", [O.NoTrans, O.Trans])
def test_refsyrk_exact(orc, uplo, trans):
    r = rng(4)
    n, k = 6, 4
    a = small_int_mat(r, n, k) if trans == O.NoTrans else small_int_mat(r, k, n)
    c = small_int_mat(r, n, n)
    cc = c.copy(order="F")
    O.refsyrk(uplo, trans, n, k, -1.0, a, 2.0, cc)
    opa = a if trans == O.NoTrans else a.T
    full = -opa @ opa.T + 2.0 * c
    tri = np.tril if uplo == O.Lower else np.triu
    other = np.triu if uplo == O.Lower else np.tril
    assert np.array_equal(tri(cc), tri(full))
    off = 1 if uplo == O.Lower else -1
    assert np.array_equal(other(cc, off), other(c, off))  # other triangle untouched


# ---------------------------------------------------------------- TRSM

@pytest.mark.parametrize("side", [O.Left, O.Right])
@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
def test_reftrsm_invariant(orc, side, uplo, ta, diag):
    r = rng(5)
    m, n = 7, 5
    na = m if side == O.Left else n
    a = np.asfortranarray(r.uniform(-1, 1, (na, na)))
    O.trddom(a, uplo, na)
    if diag == O.Unit:
        a[np.arange(na), np.arange(na)] = np.nan   # never read
    b = np.asfortranarray(r.uniform(-1, 1, (m, n)))
    x = b.copy(order="F")
    O.reftrsm(side, uplo, ta, diag, m, n, 0.7, a, x)
    assert np.all(np.isfinite(x))
    t = np.tril(a) if uplo == O.Lower else np.triu(a)
    if diag == O.Unit:
        t = t.copy()
        t[np.arange(na), np.arange(na)] = 1.0
    opa = t if ta == O.NoTrans else t.T
    lhs = opa @ x if side == O.Left else x @ opa
    assert np.allclose(lhs, 0.7 * b, rtol=0, atol=1e-13)


# ---------------------------------------------------------------- LU

def test_idamax_first_max_and_nan(orc):
    assert O.idamax([1.0, -3.0, 3.0, 2.0]) == 1
    assert O.idamax([0.0, 0.0]) == 0
    assert O.idamax([np.nan, 1.0, -2.0]) == 2
    assert O.idamax([np.nan, np.nan]) == 0


def test_getf2_pivots_match_lapack(orc):
    r = rng(6)
    for (m, n) in ((12, 12), (15, 9), (9, 15)):
        a = np.asfortranarray(r.uniform(-1, 1, (m, n)))
        lu = a.copy(order="F")
        ipiv, info = O.getf2(lu)
        assert info == 0
        mn = min(m, n)
        if m >= n:
            lu_ref, piv_ref = sla.lu_factor(a)
            assert np.array_equal(ipiv, piv_ref[:mn])
            assert np.allclose(lu, lu_ref, atol=1e-12)
        assert O.lu_resid(m, n, a, lu, ipiv) < 10.0


def test_getf2_known_pivots(orc):
    n = 8
    # permutation matrix: pivots are determined exactly
    perm = np.array([3, 0, 7, 1, 6, 2, 5, 4])
    p = np.zeros((n, n), order="F")
    p[perm, np.arange(n)] = 1.0
    lu = p.copy(order="F")
    ipiv, info = O.getf2(lu)
    assert info == 0
    assert np.allclose(lu, np.eye(n))
    # applying the recorded swaps to the identity permutation reproduces perm
    rows = np.arange(n)
    for i, q in enumerate(ipiv):
        rows[[i, q]] = rows[[q, i]]
    assert np.array_equal(p[rows], np.eye(n))
    # column-diagonally-dominant: no row moves
    r = rng(7)
    d = np.asfortranarray(r.uniform(-1, 1, (n, n)) + np.diag(np.full(n, 2.0 * n)))
    ipiv, info = O.getf2(d.copy(order="F"))
    assert list(ipiv) == list(range(n))
    # A = P^T L U with |L_ij| < 1/2: pivots are the rows of P
    L = np.tril(r.uniform(-0.45, 0.45, (n, n)), -1) + np.eye(n)
    U = np.triu(r.uniform(0.5, 1.0, (n, n)))
    order = r.permutation(n)
    a = np.asfortranarray((L @ U)[np.argsort(order)])
    ipiv, info = O.getf2(a.copy(order="F"))
    rows = np.arange(n)
    for i, q in enumerate(ipiv):
        rows[[i, q]] = rows[[q, i]]
    assert np.array_equal(np.argsort(order)[rows], np.arange(n))


def test_getf2_zero_column_info_continues(orc):
    r = rng(8)
    a = np.asfortranarray(r.uniform(-1, 1, (6, 6)))
    a[:, 2] = 0.0
    lu = a.copy(order="F")
    ipiv, info = O.getf2(lu)
    assert info == 3
    assert np.all(np.isfinite(lu))
    assert O.lu_resid(6, 6, a, lu, ipiv) < 10.0


def test_laswp_matches_numpy(orc):
    r = rng(9)
    a = np.asfortranarray(r.standard_normal((10, 4)))
    ipiv = np.array([4, 1, 9, 3, 7, 5, 8, 7, 8, 9], dtype=np.int32)
    for k1, k2, inc in ((0, 10, 1), (2, 7, 1), (0, 10, -1)):
        b = a.copy(order="F")
        O.laswp(b, k1, k2, ipiv, inc)
        rows = np.arange(10)
        rng_i = range(k1, k2) if inc > 0 else range(k2 - 1, k1 - 1, -1)
        for i in rng_i:
            rows[[i, ipiv[i]]] = rows[[ipiv[i], i]]
        assert np.array_equal(b, a[rows])


def test_npvts_counts_moved_rows(orc):
    # lutst's NPVTS: entries with ipiv[i] != i (identity pivots count zero)
    assert O.npvts(np.arange(7, dtype=np.int32)) == 0
    ipiv = np.array([4, 1, 9, 3, 7, 5, 8, 7, 8, 9], dtype=np.int32)
    assert O.npvts(ipiv) == int(np.count_nonzero(ipiv != np.arange(10)))
    assert O.npvts(ipiv) == 4


def test_lu_resid_detects_planted_error(orc):
    r = rng(10)
    a = np.asfortranarray(r.uniform(-1, 1, (20, 20)))
    lu = a.copy(order="F")
    ipiv, _ = O.getf2(lu)
    base = O.lu_resid(20, 20, a, lu, ipiv)
    assert base < 10
    bad = lu.copy(order="F")
    bad[15, 4] += 1e-9                      # a wrong multiplier
    assert O.lu_resid(20, 20, a, bad, ipiv) > 1e4
    wrong_piv = ipiv.copy()
    wrong_piv[3] = 3 if ipiv[3] != 3 else 4  # a wrong pivot record
    assert O.lu_resid(20, 20, a, lu, wrong_piv) > 1e4


# ---------------------------------------------------------------- Cholesky

@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_potf2_known_factor(orc, uplo):
    r = rng(11)
    n = 9
    L0 = np.tril(r.uniform(-1, 1, (n, n)), -1) + np.diag(r.uniform(1, 2, n))
    a = L0 @ L0.T
    s = np.asfortranarray(np.tril(a) if uplo == O.Lower else np.triu(a))
    f = s.copy(order="F")
    assert O.potf2(uplo, f) == 0
    got = np.tril(f) if uplo == O.Lower else np.triu(f).T
    assert np.allclose(got, L0, atol=1e-12)
    off = -1 if uplo == O.Lower else 1
    other = np.triu if uplo == O.Lower else np.tril
    assert np.array_equal(other(f, -off), other(s, -off))  # other triangle untouched
    assert O.llt_resid(uplo, n, s, f) < 10.0


def test_potf2_info_indefinite(orc):
    a = np.asfortranarray(np.diag([4.0, 1.0, -2.0, 5.0]))
    assert O.potf2(O.Lower, a.copy(order="F")) == 3
    b = np.asfortranarray(np.diag([4.0, np.nan, 1.0]))
    assert O.potf2(O.Lower, b) == 2


def test_llt_resid_detects_planted_error(orc):
    n = 12
    s = O.lltgen(O.Lower, n, n * 1029 + n)
    f = s.copy(order="F")
    assert O.potf2(O.Lower, f) == 0
    assert O.llt_resid(O.Lower, n, s, f) < 10
    f[7, 3] *= 1.0 + 1e-10
    assert O.llt_resid(O.Lower, n, s, f) > 1e3


# ---------------------------------------------------------------- solves

def test_getrs_potrs_against_numpy(orc):
    r = rng(12)
    n, nrhs = 11, 3
    a = np.asfortranarray(r.uniform(-1, 1, (n, n)))
    xc = np.asfortranarray(r.uniform(-1, 1, (n, nrhs)))
    b = np.asfortranarray(a @ xc)
    lu = a.copy(order="F")
    ipiv, info = O.getf2(lu)
    for trans, rhs in ((O.NoTrans, b), (O.Trans, np.asfortranarray(a.T @ xc))):
        x = rhs.copy(order="F")
        O.getrs(trans, lu, ipiv, x)
        assert np.allclose(x, xc, atol=1e-10)
    assert O.slv_resid(n, nrhs, O.genrm1(a), xc, x) < 100
    spd = np.asfortranarray(a @ a.T + n * np.eye(n))
    for uplo in (O.Lower, O.Upper):
        f = spd.copy(order="F")
        assert O.potf2(uplo, f) == 0
        x = np.asfortranarray(spd @ xc)
        O.potrs(uplo, f, x)
        assert np.allclose(x, xc, atol=1e-10)


def test_slv_resid_formula_pins(orc):
    xc = np.asfortranarray(np.array([[1.0], [-4.0], [2.0]]))
    x = xc.copy(order="F")
    x[1, 0] += 3 * 2.0 ** -52 * 4
    # ||x-xc||_inf / (N eps nrmA ||xc||_inf) = (12 eps)/(3 eps 2 4) = 0.5
    assert O.slv_resid(3, 1, 2.0, xc, x) == pytest.approx(0.5, rel=1e-12)


def test_gemm_resid_and_elem_bound(orc):
    r = rng(13)
    m, n, k = 9, 7, 5
    a = np.asfortranarray(r.uniform(-1, 1, (m, k)))
    b = np.asfortranarray(r.uniform(-1, 1, (k, n)))
    c0 = np.asfortranarray(r.uniform(-1, 1, (m, n)))
    ct = c0.copy(order="F")
    O.refgemm(O.NoTrans, O.NoTrans, m, n, k, 1.0, a, b, 1.0, ct)
    assert O.gemm_resid(O.NoTrans, O.NoTrans, m, n, k, 1.0, a, b, 1.0, c0, ct, ct) == 0.0
    cx = ct.copy(order="F")
    cx[2, 3] += 1e-6
    denom = (max(O.genrm1(a), 1) * max(O.genrm1(b), 1) * max(O.genrm1(c0), 1)
             * 2.0 ** -52 * max(m, n, k))
    got = O.gemm_resid(O.NoTrans, O.NoTrans, m, n, k, 1.0, a, b, 1.0, c0, cx, ct)
    assert got == pytest.approx(1e-6 / denom, rel=1e-6)
    nbad, _ = O.gemm_elem_check(m, n, k, 1.0, 1.0, ct, ct)
    assert nbad == 0
    nbad, mx = O.gemm_elem_check(m, n, k, 1.0, 1.0, cx, ct)
    assert nbad == 1 and mx == pytest.approx(1e-6)
    cx[0, 0] = np.nan
    assert O.gemm_elem_check(m, n, k, 1.0, 1.0, cx, ct)[0] == 2


def test_trsm_resid_formula(orc):
    r = rng(14)
    m, n = 6, 4
    a = np.asfortranarray(r.uniform(-1, 1, (m, m)))
    O.trddom(a, O.Lower, m)
    b0 = np.asfortranarray(r.uniform(-1, 1, (m, n)))
    x = b0.copy(order="F")
    O.reftrsm(O.Left, O.Lower, O.NoTrans, O.Unit, m, n, 1.0, a, x)
    assert O.trsm_resid(O.Left, O.Lower, O.Unit, m, n, 1.0, a, b0, x, x) == 0.0
    y = x.copy(order="F")
    y[1, 1] += 1e-8
    denom = max(O.trnrm1(O.Lower, O.Unit, a, m), 1) * max(O.genrm1(b0), 1) * 2.0 ** -52 * max(m, n)
    assert O.trsm_resid(O.Left, O.Lower, O.Unit, m, n, 1.0, a, b0, y, x) == pytest.approx(
        1e-8 / denom, rel=1e-6)


def test_params_fixture_present():
    with open(os.path.join(GOLD, "atlas_test_params.json")) as f:
        p = json.load(f)
    assert p["cblat3"]["N"] == [0, 1, 2, 3, 5, 9]
    assert p["thresholds"]["l3blastst"] == 50.0


# ---- TRMM, SYMM, SYR2K, TRTRI, LAUUM against numpy ------------------------

def _np_op_tri(a, n, uplo, ta, diag):
    t = np.tril(a[:n, :n]) if uplo == O.Lower else np.triu(a[:n, :n])
    if diag == O.Unit:
        np.fill_diagonal(t, 1.0)
    return t.T if ta != O.NoTrans else t


@pytest.mark.parametrize("side", [O.Left, O.Right])
@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
def test_reftrmm_against_numpy(orc, side, uplo, ta, diag):
    r = rng(20)
    m, n = 7, 5
    na = m if side == O.Left else n
    a = np.asfortranarray(r.uniform(-1, 1, (na + 2, na)))
    a[np.diag_indices(na)] = np.nan if diag == O.Unit else 2.0  # unit diagonal never read
    b0 = np.asfortranarray(r.uniform(-1, 1, (m + 1, n)))
    b = b0.copy(order="F")
    O.reftrmm(side, uplo, ta, diag, m, n, 1.5, a, b)
    t = _np_op_tri(a, na, uplo, ta, diag)
    want = 1.5 * (t @ b0[:m] if side == O.Left else b0[:m] @ t)
    assert np.allclose(b[:m], want, rtol=1e-14, atol=1e-14)
    assert np.array_equal(b[m:], b0[m:])


@pytest.mark.parametrize("side", [O.Left, O.Right])
@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_refsymm_against_numpy(orc, side, uplo):
    r = rng(21)
    m, n = 6, 4
    na = m if side == O.Left else n
    a = np.asfortranarray(r.uniform(-1, 1, (na, na)))
    s = np.tril(a) + np.tril(a, -1).T if uplo == O.Lower else np.triu(a) + np.triu(a, 1).T
    if uplo == O.Lower:  # the other triangle is never read
        a[np.triu_indices(na, 1)] = np.nan
    else:
        a[np.tril_indices(na, -1)] = np.nan
    b = np.asfortranarray(r.uniform(-1, 1, (m, n)))
    c0 = np.asfortranarray(r.uniform(-1, 1, (m, n)))
    c = c0.copy(order="F")
    O.refsymm(side, uplo, m, n, 0.7, a, b, 1.3, c)
    want = 0.7 * (s @ b if side == O.Left else b @ s) + 1.3 * c0
    assert np.allclose(c, want, rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("trans", [O.NoTrans, O.Trans])
def test_refsyr2k_against_numpy(orc, uplo, trans):
    r = rng(22)
    n, k = 6, 3
    shape = (n, k) if trans == O.NoTrans else (k, n)
    a = np.asfortranarray(r.uniform(-1, 1, shape))
    b = np.asfortranarray(r.uniform(-1, 1, shape))
    c0 = np.asfortranarray(r.uniform(-1, 1, (n, n)))
    c = c0.copy(order="F")
    O.refsyr2k(uplo, trans, n, k, 0.7, a, b, 0.0 if uplo == O.Lower else 1.3, c)
    oa, ob = (a, b) if trans == O.NoTrans else (a.T, b.T)
    beta = 0.0 if uplo == O.Lower else 1.3
    full = 0.7 * (oa @ ob.T + ob @ oa.T) + beta * c0
    mask = np.tril(np.ones((n, n), bool)) if uplo == O.Lower else np.triu(np.ones((n, n), bool))
    assert np.allclose(c[mask], full[mask], rtol=1e-14, atol=1e-14)
    assert np.array_equal(c[~mask], c0[~mask])


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
def test_reftrtri_against_numpy(orc, uplo, diag):
    r = rng(23)
    n = 9
    a = np.asfortranarray(r.uniform(-1, 1, (n, n)) / n + np.diag(r.uniform(1, 2, n)))
    a0 = a.copy(order="F")
    assert O.reftrtri(uplo, diag, a) == 0
    t = _np_op_tri(a0, n, uplo, O.NoTrans, diag)
    inv = np.linalg.inv(t)
    mask = np.tril(np.ones((n, n), bool)) if uplo == O.Lower else np.triu(np.ones((n, n), bool))
    if diag == O.Unit:
        mask &= ~np.eye(n, dtype=bool)
    assert np.allclose(a[mask], inv[mask], rtol=1e-13, atol=1e-13)
    assert np.array_equal(a[~mask], a0[~mask])
    # a zero on a NonUnit diagonal: its 1-based index, A untouched
    z = a0.copy(order="F")
    z[4, 4] = 0.0
    z0 = z.copy(order="F")
    assert O.reftrtri(uplo, O.NonUnit, z) == 5 and np.array_equal(z, z0)


@pytest.mark.parametrize("uplo", [O.Lower, O.Upper])
def test_reflauum_against_numpy(orc, uplo):
    r = rng(24)
    n = 8
    a = np.asfortranarray(r.uniform(-1, 1, (n, n)))
    a0 = a.copy(order="F")
    O.reflauum(uplo, a)
    if uplo == O.Lower:
        l = np.tril(a0)
        want, mask = l.T @ l, np.tril(np.ones((n, n), bool))
    else:
        u = np.triu(a0)
        want, mask = u @ u.T, np.triu(np.ones((n, n), bool))
    assert np.allclose(a[mask], want[mask], rtol=1e-14, atol=1e-14)
    assert np.array_equal(a[~mask], a0[~mask])


# ---- c_dblat3's machinery (DBEG, DMAKE, DMMCH) ------------------------------

def test_blat_dmmch_exact_data_self_check(orc):
    # the self-check c_dblat3 runs before its tests
    # (AtlasBase/Clint/misc/CBLATST/c_dblat3.f:212-280): AB(i,j) =
    # max(i-j+1, 0), a vector j, and the closed-form product
    # CC(j) = j((j+1)j)/2 - ((j+1)j(j-1))/3 must give ERR = 0 for all four
    # transpose pairs
    n, nmax = 32, 65
    ab = np.zeros((nmax, 2 * nmax + 1), order="F")
    for j in range(1, n + 1):
        for i in range(1, n + 1):
            ab[i - 1, j - 1] = max(i - j + 1, 0)
        ab[j - 1, nmax] = j
        ab[0, nmax + j] = j
    c = np.zeros((nmax, 1), order="F")
    a = np.asfortranarray(ab[:, :nmax])
    bcol = np.asfortranarray(ab[:, nmax:nmax + 1])          # B(k, 1) = k
    brow = np.asfortranarray(ab[:1, nmax + 1:nmax + 1 + n])  # B(1, k) = k (read as B^T)
    cc = np.zeros((nmax, 1), order="F")
    for j in range(1, n + 1):
        cc[j - 1, 0] = j * ((j + 1) * j) // 2 - ((j + 1) * j * (j - 1)) // 3
    eps = O.blat_eps()
    fatal, err, emax = O.blat_dmmch(O.NoTrans, O.NoTrans, n, 1, n, 1.0, a, bcol, 0.0, c, cc, eps)
    assert not fatal and err == 0.0 and emax == 0.0
    fatal, err, _ = O.blat_dmmch(O.NoTrans, O.Trans, n, 1, n, 1.0, a, brow, 0.0, c, cc, eps)
    assert not fatal and err == 0.0
    # reversed vector with transposed A
    for j in range(1, n + 1):
        bcol[j - 1, 0] = n - j + 1
        brow[0, j - 1] = n - j + 1
    for j in range(1, n + 1):
        cc[n - j, 0] = j * ((j + 1) * j) // 2 - ((j + 1) * j * (j - 1)) // 3
    fatal, err, _ = O.blat_dmmch(O.Trans, O.NoTrans, n, 1, n, 1.0, a, bcol, 0.0, c, cc, eps)
    assert not fatal and err == 0.0
    fatal, err, _ = O.blat_dmmch(O.Trans, O.Trans, n, 1, n, 1.0, a, brow, 0.0, c, cc, eps)
    assert not fatal and err == 0.0
    # a planted error: the ratio is |d| / (eps * gauge) for that row, and a
    # large one is fatal (err * sqrt(eps) >= 1).  Forward data, NoTrans:
    # row 4 has gauge sum_k |A(4,k)| k = 20 = its exact value.
    for j in range(1, n + 1):
        bcol[j - 1, 0] = j
        cc[j - 1, 0] = j * ((j + 1) * j) // 2 - ((j + 1) * j * (j - 1)) // 3
    assert cc[3, 0] == 20.0
    bad = cc.copy(order="F")
    bad[3, 0] += 20.0 * 4.0 * eps
    fatal, err, _ = O.blat_dmmch(O.NoTrans, O.NoTrans, n, 1, n, 1.0, a, bcol, 0.0, c, bad, eps)
    assert not fatal and err == pytest.approx(4.0, rel=1e-12)
    bad[3, 0] += 1.0
    assert O.blat_dmmch(O.NoTrans, O.NoTrans, n, 1, n, 1.0, a, bcol, 0.0, c, bad, eps)[0]


def test_blat_dbeg_and_dmake_structure(orc):
    g = O.BlatGen()
    vals = [g.dbeg() for _ in range(400)]
    # values (i - 500) / 1001 with i in [0, 999]
    ints = [round(v * 1001 + 500) for v in vals]
    assert all(0 <= i <= 999 for i in ints)
    assert all(abs((i - 500) / 1001.0 - v) < 1e-15 for i, v in zip(ints, vals))
    # the sequence: i <- 891 i mod 1000 from 7, every fifth step skipped
    i, ic, want = 7, 0, []
    while len(want) < 400:
        ic += 1
        i = (i * 891) % 1000
        if ic >= 5:
            ic = 0
            i = (i * 891) % 1000
        want.append(i)
    assert ints == want
    g.reset()
    assert g.dbeg() == vals[0]
    # DMAKE: symmetric copy, zero column n/2 off the diagonal, triangles
    # zeroed, +1 on a triangular diagonal, rogue padding
    n, nmax, lda = 6, 8, 7
    g.reset()
    s, ss = O.blat_dmake(g, O.BLAT_SY, O.Lower, O.NonUnit, n, n, nmax, lda)
    assert np.array_equal(s[:n, :n], s[:n, :n].T)
    col = n // 2 - 1
    assert np.all(s[col + 1:n, col] == 0.0)  # the generated part of column n/2
    assert np.all(ss[n:lda, :n] == -1.0e10)
    assert np.all(ss[np.triu_indices(n, 1)] == -1.0e10)          # upper part is rogue
    g.reset()
    t, tt = O.blat_dmake(g, O.BLAT_TR, O.Upper, O.Unit, n, n, nmax, lda)
    assert np.all(np.diag(t[:n, :n]) == 1.0)
    assert np.all(t[:n, :n][np.tril_indices(n, -1)] == 0.0)
    assert np.all(np.diag(tt[:n, :n]) == -1.0e10)                # unit diagonal is rogue
    g.reset()
    t, _ = O.blat_dmake(g, O.BLAT_TR, O.Lower, O.NonUnit, n, n, nmax, lda)
    g.reset()
    gen = O.BlatGen()
    first = gen.dbeg()
    assert t[0, 0] == pytest.approx(first + 1.0)
