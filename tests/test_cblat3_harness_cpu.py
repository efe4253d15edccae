"""The c_dblat3 restatement's own harness, checked on the CPU: driven with
a numpy stand-in for the C ABI it passes, and each of three planted
mistakes in the stand-in (a dropped SYR2K term, a unit diagonal ignored by
TRMM, one corrupted GEMM element) makes it fail.  This pins the harness
(argument mapping for both storage orders, the DMMCH calls, the padding
checks) independently of the GPU library."""
import numpy as np
import pytest

import oracle as O
import test_gpu_cblat3 as T

CM, RM = 101, 102


def _get(arr, order, r, c, ld):
    m = arr.reshape(-1, order="F")
    i, j = np.meshgrid(np.arange(r), np.arange(c), indexing="ij")
    return m[i + j * ld] if order == CM else m[i * ld + j]


def _put(arr, order, M, mask, ld):
    m = arr.reshape(-1, order="F")  # a view: arr is Fortran-contiguous
    i, j = np.nonzero(mask)
    m[(i + j * ld) if order == CM else (i * ld + j)] = M[i, j]


def _op(M, t):
    return M if t == O.NoTrans else M.T


def _tri(A, uplo, diag):
    t = np.triu(A) if uplo == O.Upper else np.tril(A)
    if diag == O.Unit:
        np.fill_diagonal(t, 1.0)
    return t


def _sym(A, uplo):
    t = np.triu(A) if uplo == O.Upper else np.tril(A)
    return t + t.T - np.diag(np.diag(t))


class Fake:
    ColMajor, RowMajor = CM, RM
    drop_syr2k_term = False
    ignore_unit = False
    corrupt_gemm = False

    def cblas_dgemm(self, order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        A = _get(a, order, *((m, k) if ta == O.NoTrans else (k, m)), lda)
        B = _get(b, order, *((k, n) if tb == O.NoTrans else (n, k)), ldb)
        C = _get(c, order, m, n, ldc)
        R = alpha * _op(A, ta) @ _op(B, tb) + (beta * C if beta != 0 else 0)
        if self.corrupt_gemm and m > 1 and n > 1:
            R[1, 1] += 1e-3
        _put(c, order, R, np.ones((m, n), bool), ldc)

    def cblas_dsymm(self, order, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
        na = m if side == O.Left else n
        A = _sym(_get(a, order, na, na, lda), uplo)
        B, C = _get(b, order, m, n, ldb), _get(c, order, m, n, ldc)
        R = alpha * (A @ B if side == O.Left else B @ A) + (beta * C if beta != 0 else 0)
        _put(c, order, R, np.ones((m, n), bool), ldc)

    def _trx(self, solve, order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
        na = m if side == O.Left else n
        d = O.NonUnit if self.ignore_unit else diag
        A = _op(_tri(_get(a, order, na, na, lda), uplo, d), ta)
        B = _get(b, order, m, n, ldb)
        if solve:
            R = alpha * (np.linalg.solve(A, B) if side == O.Left else np.linalg.solve(A.T, B.T).T)
        else:
            R = alpha * (A @ B if side == O.Left else B @ A)
        _put(b, order, R, np.ones((m, n), bool), ldb)

    def cblas_dtrmm(self, *a):
        self._trx(False, *a)

    def cblas_dtrsm(self, *a):
        self._trx(True, *a)

    def _tri_mask(self, n, uplo):
        return np.triu(np.ones((n, n), bool)) if uplo == O.Upper else np.tril(np.ones((n, n), bool))

    def cblas_dsyrk(self, order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
        A = _op(_get(a, order, *((n, k) if trans == O.NoTrans else (k, n)), lda), trans)
        C = _get(c, order, n, n, ldc)
        R = alpha * A @ A.T + (beta * C if beta != 0 else 0)
        _put(c, order, R, self._tri_mask(n, uplo), ldc)

    def cblas_dsyr2k(self, order, uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
        shp = (n, k) if trans == O.NoTrans else (k, n)
        A, B = _op(_get(a, order, *shp, lda), trans), _op(_get(b, order, *shp, ldb), trans)
        C = _get(c, order, n, n, ldc)
        S = A @ B.T + (A @ B.T if self.drop_syr2k_term else B @ A.T)
        R = alpha * S + (beta * C if beta != 0 else 0)
        _put(c, order, R, self._tri_mask(n, uplo), ldc)


@pytest.fixture
def small(monkeypatch):
    monkeypatch.setattr(T, "NS", [0, 1, 2, 5])
    monkeypatch.setattr(T, "ALF", [0.0, 0.7])
    monkeypatch.setattr(T, "BET", [1.3])


def _all(F, order):
    T.test_cblat3_dgemm(F, order)
    T.test_cblat3_dsymm(F, order)
    for r in ("trmm", "trsm"):
        T.test_cblat3_dtrmm_dtrsm(F, r, order)
    for r in ("syrk", "syr2k"):
        T.test_cblat3_dsyrk_dsyr2k(F, r, order)


@pytest.mark.parametrize("order", ["col", "row"])
def test_harness_passes_a_correct_library(small, order):
    _all(Fake(), order)


@pytest.mark.parametrize("flaw,run", [
    ("drop_syr2k_term", lambda F: T.test_cblat3_dsyrk_dsyr2k(F, "syr2k", "row")),
    ("ignore_unit", lambda F: T.test_cblat3_dtrmm_dtrsm(F, "trmm", "col")),
    ("corrupt_gemm", lambda F: T.test_cblat3_dgemm(F, "col")),
])
def test_harness_catches_planted_mistakes(small, flaw, run):
    F = Fake()
    setattr(F, flaw, True)
    with pytest.raises(AssertionError):
        run(F)
