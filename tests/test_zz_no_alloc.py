"""atlg_dtrsm_noalloc (the allocation-free TRSM entry point) against
atlg_dtrsm, through the C ABI: two calls in a row, the second smaller, each
bit-identical to atlg_dtrsm on the same inputs (both take the blocked solve
there); a few-RHS case (where atlg_dtrsm takes its sync-free solve) within
l3blastst's residual bound; and orders above the maximum rejected with
ATLG_ESIZE, leaving B untouched."""
import ctypes

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


def _tri(rng, na, uplo):
    # trddom-style well-conditioned triangle: off-diagonal entries / na,
    # diagonal in [1, 2]
    a = rng.uniform(-1, 1, (na, na)) / na + np.diag(rng.uniform(1, 2, na))
    a = np.tril(a) if uplo == O.Lower else np.triu(a)
    return np.asfortranarray(a)


def _run(G, fn, side, uplo, ta, diag, m, n, a, b0):
    import torch
    L = G.lib()
    na = m if side == G.Left else n
    da = G.from_fortran(a)
    db = G.from_fortran(b0)
    st = torch.cuda.current_stream()
    rc = getattr(L, fn)(side, uplo, ta, diag, m, n, 1.5, ctypes.c_void_p(da.data_ptr()), na,
                        ctypes.c_void_p(db.data_ptr()), m, ctypes.c_void_p(st.cuda_stream))
    torch.cuda.synchronize()
    return rc, db


def test_noalloc_matches_dtrsm(G):
    import torch
    rng = np.random.default_rng(7)
    cases = [(G.Left, G.Lower, G.NoTrans, G.NonUnit, 3000, 37),   # first call
             (G.Right, G.Upper, G.Trans, G.Unit, 700, 300)]       # second, smaller
    for side, uplo, ta, diag, m, n in cases:
        na = m if side == G.Left else n
        a = _tri(rng, na, uplo)
        b0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
        rc1, ref = _run(G, "atlg_dtrsm", side, uplo, ta, diag, m, n, a, b0)
        rc2, got = _run(G, "atlg_dtrsm_noalloc", side, uplo, ta, diag, m, n, a, b0)
        assert rc1 == 0 and rc2 == 0
        assert torch.equal(got, ref), (side, uplo, ta, diag, m, n)
        # and both solve the system (textbook TRSM of the oracle)
        tr = b0.copy(order="F")
        O.reftrsm(side, uplo, ta, diag, m, n, 1.5, a, tr)
        r = O.trsm_resid(side, uplo, diag, m, n, 1.5, a, b0, G.to_fortran(got), tr)
        assert r < 50.0, r


def test_noalloc_few_rhs_and_limit(G):
    import torch
    rng = np.random.default_rng(8)
    m, n = 2048, 2
    a = _tri(rng, m, G.Lower)
    b0 = np.asfortranarray(rng.uniform(-1, 1, (m, n)))
    _, ref = _run(G, "atlg_dtrsm", G.Left, G.Lower, G.NoTrans, G.NonUnit, m, n, a, b0)
    rc, got = _run(G, "atlg_dtrsm_noalloc", G.Left, G.Lower, G.NoTrans, G.NonUnit, m, n, a, b0)
    assert rc == 0
    r = O.trsm_resid(G.Left, G.Lower, G.NonUnit, m, n, 1.5, a, b0, G.to_fortran(got),
                     G.to_fortran(ref))
    assert r < 50.0, r
    # order 8193: rejected, nothing queued
    m = 8193
    a = np.asfortranarray(np.eye(m))
    b0 = np.asfortranarray(rng.uniform(-1, 1, (m, 1)))
    rc, got = _run(G, "atlg_dtrsm_noalloc", G.Left, G.Lower, G.NoTrans, G.NonUnit, m, 1, a, b0)
    assert rc == G.ESIZE
    assert torch.equal(got, G.from_fortran(b0))
    assert b"8192" in G.lib().atlg_last_error()
