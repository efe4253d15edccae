"""GPU parity for the level-3 neighbours of the hot path (TRMM, SYMM,
SYR2K), the inverses built on them (TRTRI, LAUUM, POTRI) and the Fortran 77
entry points, against the oracle's restatements of ATLAS's reference
kernels (AtlasBase/Clint/atlas-tst.base ATL_dreftrmm / ATL_drefsymm /
ATL_drefsyr2k; ATL_dtrtri / ATL_dlauum recursions in atlas-lp.base) at
sizes that exercise the recursions (orders above the 128 leaf, ragged
ends), checked with c_dblat3's DMMCH gauge where the result is a product.
"""
import ctypes

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

THRESH = 16.0


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def _rng(seed):
    return np.random.default_rng(seed)


def _f(x):
    return np.asfortranarray(x, dtype=np.float64)


def _tri(a, uplo, diag):
    t = np.triu(a) if uplo == O.Upper else np.tril(a)
    if diag == O.Unit:
        np.fill_diagonal(t, 1.0)
    return t


def _sym(a, uplo):
    t = np.triu(a) if uplo == O.Upper else np.tril(a)
    return t + t.T - np.diag(np.diag(a))


def _mmch(ta, tb, m, n, k, alpha, a, b, beta, c, cc):
    fatal, _, emax = O.blat_dmmch(ta, tb, m, n, k, alpha, _f(a), _f(b), beta, _f(c), _f(cc))
    assert not fatal and emax < THRESH, emax
    return emax


def _ldpad(a, pad=3):
    """a stored with leading dimension rows + pad (rogue padding)."""
    r, c = a.shape
    x = np.full((r + pad, c), -1.0e10, order="F")
    x[:r] = a
    return x


@pytest.mark.parametrize("side", [O.Left, O.Right])
@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("ta", [O.NoTrans, O.Trans])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
def test_trmm_recursive(G, side, uplo, ta, diag):
    rng = _rng(side * 7 + uplo * 3 + ta + diag)
    m, n = 517, 389
    na = m if side == O.Left else n
    a = rng.uniform(-1, 1, (na, na))
    b = rng.uniform(-1, 1, (m, n))
    # the other triangle (and a unit diagonal) holds garbage that must be ignored
    araw = a.copy()
    araw[(np.tril(np.ones((na, na)), -1) if uplo == O.Upper else np.triu(np.ones((na, na)), 1))
         .astype(bool)] = 1.0e3
    if diag == O.Unit:
        np.fill_diagonal(araw, 7.0)
    A, B = _ldpad(araw), _ldpad(b)
    G.cblas_dtrmm(G.ColMajor, side, uplo, ta, diag, m, n, 0.7, A, A.shape[0], B, B.shape[0])
    assert np.all(B[m:] == -1.0e10)
    t = _tri(araw, uplo, diag)
    zero = np.zeros((m, n), order="F")
    if side == O.Left:
        _mmch(ta, O.NoTrans, m, n, m, 0.7, t, b, 0.0, zero, B[:m])
    else:
        _mmch(O.NoTrans, ta, m, n, n, 0.7, b, t, 0.0, zero, B[:m])
    # the oracle's ATL_dreftrmm agrees too
    ref = _f(b.copy())
    O.reftrmm(side, uplo, ta, diag, m, n, 0.7, _f(araw), ref)
    np.testing.assert_allclose(B[:m], ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("side", [O.Left, O.Right])
@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
def test_symm(G, side, uplo):
    rng = _rng(11 + side + uplo)
    m, n = 300, 451
    na = m if side == O.Left else n
    a = rng.uniform(-1, 1, (na, na))
    b, c = rng.uniform(-1, 1, (m, n)), rng.uniform(-1, 1, (m, n))
    s = _sym(a, uplo)
    araw = a.copy()  # the other triangle differs from the mirrored one
    A, B, C = _ldpad(araw), _ldpad(b), _ldpad(c)
    G.cblas_dsymm(G.ColMajor, side, uplo, m, n, -0.4, A, A.shape[0], B, B.shape[0], 1.3, C,
                  C.shape[0])
    assert np.all(C[m:] == -1.0e10)
    if side == O.Left:
        _mmch(O.NoTrans, O.NoTrans, m, n, m, -0.4, s, b, 1.3, c, C[:m])
    else:
        _mmch(O.NoTrans, O.NoTrans, m, n, n, -0.4, b, s, 1.3, c, C[:m])
    ref = _f(c.copy())
    O.refsymm(side, uplo, m, n, -0.4, _f(araw), _f(b), 1.3, ref)
    np.testing.assert_allclose(C[:m], ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("trans", [O.NoTrans, O.Trans])
def test_syr2k(G, uplo, trans):
    rng = _rng(21 + uplo + trans)
    n, k = 403, 260
    shp = (n, k) if trans == O.NoTrans else (k, n)
    a, b = rng.uniform(-1, 1, shp), rng.uniform(-1, 1, shp)
    c = rng.uniform(-1, 1, (n, n))
    A, B, C = _ldpad(a), _ldpad(b), _ldpad(c)
    G.cblas_dsyr2k(G.ColMajor, uplo, trans, n, k, 0.9, A, A.shape[0], B, B.shape[0], 0.5, C,
                   C.shape[0])
    other = (np.tril(np.ones((n, n)), -1) if uplo == O.Upper else np.triu(np.ones((n, n)), 1))
    assert np.array_equal(C[:n][other.astype(bool)], c[other.astype(bool)])
    opa = a if trans == O.NoTrans else a.T
    opb = b if trans == O.NoTrans else b.T
    full = 0.9 * (opa @ opb.T + opb @ opa.T) + 0.5 * c
    mask = ~other.astype(bool)
    gauge = 0.9 * 2 * (np.abs(opa) @ np.abs(opb).T) + 0.5 * np.abs(c)
    ratio = np.abs(full - C[:n])[mask] / (gauge[mask] * 2.0 ** -52)
    assert ratio.max() < THRESH * 2 * k ** 0.5 + THRESH
    ref = _f(c.copy())
    O.refsyr2k(uplo, trans, n, k, 0.9, _f(a), _f(b), 0.5, ref)
    np.testing.assert_allclose(C[:n], ref, rtol=0, atol=1e-11 * np.abs(ref).max())


def _well_tri(rng, n, uplo):
    a = rng.uniform(-1, 1, (n, n)) / np.sqrt(n)
    np.fill_diagonal(a, rng.uniform(1.0, 2.0, n) * np.sign(rng.uniform(-1, 1, n)))
    return a


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
@pytest.mark.parametrize("n", [1, 127, 129, 700])
def test_trtri(G, uplo, diag, n):
    rng = _rng(n + uplo + diag)
    a = _well_tri(rng, n, uplo)
    A = _ldpad(a)
    info = G.clapack_dtrtri(G.ColMajor, uplo, diag, n, A, A.shape[0])
    assert info == 0
    t = _tri(a, uplo, diag)
    inv = _tri(A[:n], uplo, diag)
    other = (np.tril(np.ones((n, n)), -1) if uplo == O.Upper else np.triu(np.ones((n, n)), 1))
    assert np.array_equal(A[:n][other.astype(bool)], a[other.astype(bool)])
    if diag == O.Unit:
        assert np.array_equal(np.diag(A[:n]), np.diag(a))  # a unit diagonal is not written
    resid = np.abs(inv @ t - np.eye(n)).max()
    assert resid < 64 * n * 2.0 ** -52 * np.abs(inv).max() * np.abs(t).max() + 1e-15
    ref = _f(a.copy())
    assert O.reftrtri(uplo, diag, ref) == 0
    np.testing.assert_allclose(A[:n], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
def test_trtri_singular_leaves_a(G, uplo):
    rng = _rng(5)
    n = 300
    a = _well_tri(rng, n, uplo)
    a[211, 211] = 0.0
    A = _ldpad(a)
    before = A.copy()
    assert G.clapack_dtrtri(G.ColMajor, uplo, O.NonUnit, n, A, A.shape[0]) == 212
    assert np.array_equal(A, before)
    # a unit diagonal is never singular
    assert G.clapack_dtrtri(G.ColMajor, uplo, O.Unit, n, A, A.shape[0]) == 0


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("n", [1, 128, 131, 650])
def test_lauum(G, uplo, n):
    rng = _rng(40 + n + uplo)
    a = rng.uniform(-1, 1, (n, n))
    A = _ldpad(a)
    G.clapack_dlauum(G.ColMajor, uplo, n, A, A.shape[0])
    t = _tri(a, uplo, O.NonUnit)
    full = t @ t.T if uplo == O.Upper else t.T @ t
    tri = np.triu(np.ones((n, n), bool)) if uplo == O.Upper else np.tril(np.ones((n, n), bool))
    assert np.array_equal(A[:n][~tri], a[~tri])
    gauge = (np.abs(t) @ np.abs(t).T) if uplo == O.Upper else (np.abs(t).T @ np.abs(t))
    ratio = np.abs(A[:n] - full)[tri] / (gauge[tri] * 2.0 ** -52)
    assert ratio.max() < THRESH * max(1.0, np.sqrt(n))
    ref = _f(a.copy())
    O.reflauum(uplo, ref)
    np.testing.assert_allclose(A[:n], ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("diag", [O.NonUnit, O.Unit])
@pytest.mark.parametrize("n", [1024, 1500, 2300])
def test_trtri_doubling_orders(G, uplo, diag, n):
    # orders that take several doubling levels, with ragged last blocks at
    # some of them (1500: 1500-1024 = 476 rows at level 2048; 2300), both
    # triangles and diagonals: residual of inv(T) T, the other triangle
    # and a unit diagonal untouched, agreement with the textbook inverse
    rng = _rng(n + 7 * uplo + diag)
    a = _well_tri(rng, n, uplo)
    A = _ldpad(a)
    assert G.clapack_dtrtri(G.ColMajor, uplo, diag, n, A, A.shape[0]) == 0
    t = _tri(a, uplo, diag)
    inv = _tri(A[:n], uplo, diag)
    other = (np.tril(np.ones((n, n)), -1) if uplo == O.Upper else np.triu(np.ones((n, n)), 1))
    assert np.array_equal(A[:n][other.astype(bool)], a[other.astype(bool)])
    assert np.all(A[n:] == -1.0e10)
    if diag == O.Unit:
        assert np.array_equal(np.diag(A[:n]), np.diag(a))
    resid = np.abs(inv @ t - np.eye(n)).max()
    assert resid < 64 * n * 2.0 ** -52 * np.abs(inv).max() * np.abs(t).max() + 1e-15
    ref = np.linalg.inv(t)
    np.testing.assert_allclose(inv, ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("n", [1100, 2048])
def test_lauum_masked_product_orders(G, uplo, n):
    rng = _rng(90 + n + uplo)
    a = rng.uniform(-1, 1, (n, n))
    A = _ldpad(a)
    G.clapack_dlauum(G.ColMajor, uplo, n, A, A.shape[0])
    t = _tri(a, uplo, O.NonUnit)
    full = t @ t.T if uplo == O.Upper else t.T @ t
    tri = np.triu(np.ones((n, n), bool)) if uplo == O.Upper else np.tril(np.ones((n, n), bool))
    assert np.array_equal(A[:n][~tri], a[~tri])
    assert np.all(A[n:] == -1.0e10)
    gauge = (np.abs(t) @ np.abs(t).T) if uplo == O.Upper else (np.abs(t).T @ np.abs(t))
    ratio = np.abs(A[:n] - full)[tri] / (gauge[tri] * 2.0 ** -52)
    assert ratio.max() < THRESH * max(1.0, np.sqrt(n))


def _spd(rng, n):
    m = rng.uniform(-1, 1, (n, n))
    return m @ m.T / n + np.eye(n)


@pytest.mark.parametrize("uplo", [O.Upper, O.Lower])
@pytest.mark.parametrize("n", [5, 128, 900, 2100])
def test_potri(G, uplo, n):
    rng = _rng(60 + n + uplo)
    s = _spd(rng, n)
    A = _ldpad(s)
    assert G.clapack_dpotrf(G.ColMajor, uplo, n, A, A.shape[0]) == 0
    assert G.clapack_dpotri(G.ColMajor, uplo, n, A, A.shape[0]) == 0
    inv = _sym(A[:n], uplo)
    resid = np.abs(inv @ s - np.eye(n)).max()
    assert resid < 1e-12 * n, resid
    np.testing.assert_allclose(inv, np.linalg.inv(s), rtol=0,
                               atol=1e-12 * np.abs(inv).max() * np.linalg.cond(s))


def test_potri_info(G):
    n = 40
    a = np.eye(n, order="F")
    a[17, 17] = 0.0  # a zero diagonal of the factor: info 18, nothing written
    A = _ldpad(a)
    before = A.copy()
    assert G.clapack_dpotri(G.ColMajor, O.Lower, n, A, A.shape[0]) == 18
    assert np.array_equal(A, before)


def test_rowmajor_forms(G):
    rng = _rng(70)
    n = 260
    a = _well_tri(rng, n, O.Lower)
    # row-major lower is column-major upper of the transpose
    R = np.array(a, order="C")  # row-major storage of a copy of a
    Rf = np.asfortranarray(R.reshape(-1).reshape(n, n, order="F"))  # same memory
    assert G.clapack_dtrtri(G.RowMajor, O.Lower, O.NonUnit, n, Rf, n) == 0
    got = Rf.reshape(-1, order="F").reshape(n, n)  # back to the row-major view
    inv = np.tril(got)
    assert np.abs(inv @ np.tril(a) - np.eye(n)).max() < 1e-12 * n


def test_device_api_forms(G):
    import torch
    rng = _rng(80)
    n, m = 333, 270
    s = _spd(rng, n)
    b = rng.uniform(-1, 1, (n, m))
    dA = G.from_fortran(s)
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    G.dpotrf(O.Lower, n, dA, n, info)
    G.dpotri(O.Lower, n, dA, n, info)
    torch.cuda.synchronize()
    assert int(info.item()) == 0
    inv = _sym(G.to_fortran(dA), O.Lower)
    assert np.abs(inv @ s - np.eye(n)).max() < 1e-12 * n
    # device TRMM / SYMM / SYR2K / LAUUM / TRTRI against the host entry points
    t = _well_tri(rng, n, O.Upper)
    for fn_dev, fn_host in [
        (lambda A, B: G.dtrmm(O.Left, O.Upper, O.Trans, O.NonUnit, n, m, 1.1, A, n, B, n),
         lambda A, B: G.cblas_dtrmm(G.ColMajor, O.Left, O.Upper, O.Trans, O.NonUnit, n, m, 1.1,
                                    A, n, B, n)),
        (lambda A, B: G.dsymm(O.Left, O.Upper, n, m, 1.1, A, n, B, n, 0.0, B2, n),
         lambda A, B: G.cblas_dsymm(G.ColMajor, O.Left, O.Upper, n, m, 1.1, A, n, B, n, 0.0,
                                    H2, n)),
    ]:
        B2 = G.from_fortran(np.zeros((n, m)))
        H2 = np.zeros((n, m), order="F")
        dA, dB = G.from_fortran(t), G.from_fortran(b)
        hA, hB = _f(t.copy()), _f(b.copy())
        fn_dev(dA, dB)
        fn_host(hA, hB)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(G.to_fortran(dB), hB)
        np.testing.assert_array_equal(G.to_fortran(B2), H2)
    dA = G.from_fortran(t)
    G.dtrtri(O.Upper, O.NonUnit, n, dA, n, info)
    hA = _f(t.copy())
    assert G.clapack_dtrtri(G.ColMajor, O.Upper, O.NonUnit, n, hA, n) == 0
    torch.cuda.synchronize()
    np.testing.assert_array_equal(G.to_fortran(dA), hA)
    dA = G.from_fortran(t)
    G.dlauum(O.Upper, n, dA, n)
    hA = _f(t.copy())
    G.clapack_dlauum(G.ColMajor, O.Upper, n, hA, n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(G.to_fortran(dA), hA)
    a2, b2 = rng.uniform(-1, 1, (n, 100)), rng.uniform(-1, 1, (n, 100))
    c0 = _sym(rng.uniform(-1, 1, (n, n)), O.Lower)
    dC = G.from_fortran(c0)
    G.dsyr2k(O.Lower, O.NoTrans, n, 100, 0.3, G.from_fortran(a2), n, G.from_fortran(b2), n, 2.0,
             dC, n)
    hC = _f(c0.copy())
    G.cblas_dsyr2k(G.ColMajor, O.Lower, O.NoTrans, n, 100, 0.3, _f(a2), n, _f(b2), n, 2.0, hC, n)
    torch.cuda.synchronize()
    np.testing.assert_array_equal(G.to_fortran(dC), hC)


# ------------------------------------------------------------- Fortran 77

def test_f77_blas(G):
    F = G.F77
    rng = _rng(90)
    m, n, k = 150, 133, 97
    a, b, c = rng.uniform(-1, 1, (k, m)), rng.uniform(-1, 1, (k, n)), rng.uniform(-1, 1, (m, n))
    C = _f(c.copy())
    F.call("dgemm_", ("c", "t"), ("c", "N"), ("i", m), ("i", n), ("i", k), ("d", 0.5),
           ("a", _f(a)), ("i", k), ("a", _f(b)), ("i", k), ("d", -1.0), ("a", C), ("i", m))
    _mmch(O.Trans, O.NoTrans, m, n, k, 0.5, a, b, -1.0, c, C)
    t = _well_tri(rng, m, O.Lower)
    X = _f(c.copy())
    F.call("dtrsm_", ("c", "l"), ("c", "L"), ("c", "n"), ("c", "N"), ("i", m), ("i", n),
           ("d", 2.0), ("a", _f(t)), ("i", m), ("a", X), ("i", m))
    H = _f(c.copy())
    G.cblas_dtrsm(G.ColMajor, O.Left, O.Lower, O.NoTrans, O.NonUnit, m, n, 2.0, _f(t), m, H, m)
    np.testing.assert_array_equal(X, H)
    u = _f(rng.uniform(-1, 1, (n, n)))
    Y, Hy = _f(c.copy()), _f(c.copy())
    F.call("dtrmm_", ("c", "R"), ("c", "u"), ("c", "T"), ("c", "u"), ("i", m), ("i", n),
           ("d", 1.5), ("a", u), ("i", n), ("a", Y), ("i", m))
    G.cblas_dtrmm(G.ColMajor, O.Right, O.Upper, O.Trans, O.Unit, m, n, 1.5, u, n, Hy, m)
    np.testing.assert_array_equal(Y, Hy)
    S, bb = _f(rng.uniform(-1, 1, (m, m))), _f(rng.uniform(-1, 1, (m, n)))
    Z, Hz = _f(c.copy()), _f(c.copy())
    F.call("dsymm_", ("c", "L"), ("c", "l"), ("i", m), ("i", n), ("d", 1.0), ("a", S), ("i", m),
           ("a", bb), ("i", m), ("d", 0.0), ("a", Z), ("i", m))
    G.cblas_dsymm(G.ColMajor, O.Left, O.Lower, m, n, 1.0, S, m, bb, m, 0.0, Hz, m)
    np.testing.assert_array_equal(Z, Hz)
    W = _f(rng.uniform(-1, 1, (m, m)))
    Hw = W.copy(order="F")
    F.call("dsyrk_", ("c", "U"), ("c", "T"), ("i", m), ("i", k), ("d", 1.0), ("a", _f(a)),
           ("i", k), ("d", 1.0), ("a", W), ("i", m))
    G.cblas_dsyrk(G.ColMajor, O.Upper, O.Trans, m, k, 1.0, _f(a), k, 1.0, Hw, m)
    np.testing.assert_array_equal(W, Hw)
    V = _f(rng.uniform(-1, 1, (m, m)))
    Hv = V.copy(order="F")
    F.call("dsyr2k_", ("c", "l"), ("c", "n"), ("i", m), ("i", n), ("d", 1.0), ("a", C),
           ("i", m), ("a", X), ("i", m), ("d", 0.25), ("a", V), ("i", m))
    G.cblas_dsyr2k(G.ColMajor, O.Lower, O.NoTrans, m, n, 1.0, C, m, X, m, 0.25, Hv, m)
    np.testing.assert_array_equal(V, Hv)


def test_f77_lapack(G):
    F = G.F77
    rng = _rng(91)
    n, nrhs = 500, 7
    a = rng.uniform(-1, 1, (n, n))
    b = rng.uniform(-1, 1, (n, nrhs))
    A, ip = _f(a.copy()), np.zeros(n, np.int32)
    info = F.call("dgetrf_", ("i", n), ("i", n), ("a", A), ("i", n), ("a", ip), ("info", None))
    assert info == 0
    A2, ip2 = _f(a.copy()), np.zeros(n, np.int32)
    assert G.clapack_dgetrf(G.ColMajor, n, n, A2, n, ip2) == 0
    np.testing.assert_array_equal(A, A2)
    np.testing.assert_array_equal(ip, ip2 + 1)  # 1-based pivots
    assert ip.min() >= 1 and ip.max() <= n
    X = _f(b.copy())
    assert F.call("dgetrs_", ("c", "N"), ("i", n), ("i", nrhs), ("a", A), ("i", n), ("a", ip),
                  ("a", X), ("i", n), ("info", None)) == 0
    assert np.abs(a @ X - b).max() < 1e-9
    A3, ip3, X3 = _f(a.copy()), np.zeros(n, np.int32), _f(b.copy())
    assert F.call("dgesv_", ("i", n), ("i", nrhs), ("a", A3), ("i", n), ("a", ip3), ("a", X3),
                  ("i", n), ("info", None)) == 0
    np.testing.assert_array_equal(X3, X)
    # DLASWP with the 1-based pivots reapplies the row interchanges
    P = _f(a.copy())
    F.call("dlaswp_", ("i", n), ("a", P), ("i", n), ("i", 1), ("i", n), ("a", ip), ("i", 1))
    perm = np.arange(n)
    for i in range(n):
        perm[[i, ip[i] - 1]] = perm[[ip[i] - 1, i]]
    np.testing.assert_array_equal(P, a[perm])
    F.call("dlaswp_", ("i", n), ("a", P), ("i", n), ("i", 1), ("i", n), ("a", ip), ("i", -1))
    np.testing.assert_array_equal(P, a)  # INCX < 0 undoes it
    s = _spd(rng, n)
    for u in ("U", "l"):
        S = _f(s.copy())
        assert F.call("dpotrf_", ("c", u), ("i", n), ("a", S), ("i", n), ("info", None)) == 0
        Xs = _f(b.copy())
        assert F.call("dpotrs_", ("c", u), ("i", n), ("i", nrhs), ("a", S), ("i", n), ("a", Xs),
                      ("i", n), ("info", None)) == 0
        assert np.abs(s @ Xs - b).max() < 1e-10
        S2, X2 = _f(s.copy()), _f(b.copy())
        assert F.call("dposv_", ("c", u), ("i", n), ("i", nrhs), ("a", S2), ("i", n), ("a", X2),
                      ("i", n), ("info", None)) == 0
        np.testing.assert_array_equal(X2, Xs)
        assert F.call("dpotri_", ("c", u), ("i", n), ("a", S), ("i", n), ("info", None)) == 0
        inv = _sym(S, O.Upper if u == "U" else O.Lower)
        assert np.abs(inv @ s - np.eye(n)).max() < 1e-10
    T = _f(_well_tri(rng, n, O.Upper))
    T2 = T.copy(order="F")
    assert F.call("dtrtri_", ("c", "U"), ("c", "N"), ("i", n), ("a", T), ("i", n),
                  ("info", None)) == 0
    assert G.clapack_dtrtri(G.ColMajor, O.Upper, O.NonUnit, n, T2, n) == 0
    np.testing.assert_array_equal(T, T2)
    L = _f(rng.uniform(-1, 1, (n, n)))
    L2 = L.copy(order="F")
    assert F.call("dlauum_", ("c", "L"), ("i", n), ("a", L), ("i", n), ("info", None)) == 0
    G.clapack_dlauum(G.ColMajor, O.Lower, n, L2, n)
    np.testing.assert_array_equal(L, L2)


def test_f77_xerbla(G):
    F = G.F77
    one = _f(np.ones((4, 4)))
    with G.XerblaTrap() as tr:
        F.call("dgemm_", ("c", "X"), ("c", "N"), ("i", 4), ("i", 4), ("i", 4), ("d", 1.0),
               ("a", one), ("i", 4), ("a", one), ("i", 4), ("d", 0.0), ("a", one), ("i", 4))
        F.call("dtrsm_", ("c", "L"), ("c", "U"), ("c", "N"), ("c", "N"), ("i", 4), ("i", -1),
               ("d", 1.0), ("a", one), ("i", 4), ("a", one), ("i", 4))
        info = F.call("dgetrf_", ("i", 4), ("i", 4), ("a", one), ("i", 3),
                      ("a", np.zeros(4, np.int32)), ("info", None))
        assert info == -4
        info = F.call("dpotri_", ("c", "Q"), ("i", 4), ("a", one), ("i", 4), ("info", None))
        assert info == -1
    names = [(p, r.strip()) for p, r, _ in tr.calls]
    assert names == [(1, "DGEMM"), (6, "DTRSM"), (4, "DGETRF"), (1, "DPOTRI")], tr.calls
    assert "parameter number  1 had an illegal value" in tr.calls[0][2]
    np.testing.assert_array_equal(one, np.ones((4, 4)))
