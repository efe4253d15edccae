"""CPU-only checks of the C ABI: the library loads, exports every entry point
declared in include/*.h, and its argument checking (which runs before any
device work) reports the same parameter numbers as ATLAS's wrappers
(AtlasBase/Clint/cblas.base, AtlasBase/Clint/atlas-clp.base)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def G():
    import atlas_gpu
    atlas_gpu.lib()
    return atlas_gpu


def declared_functions():
    names = set()
    for fn in os.listdir(INC):
        if not fn.endswith(".h"):
            continue
        txt = open(os.path.join(INC, fn)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"\b(?:void|int|const char\s*\*)\s*\**\s*(\w+)\s*\(", txt):
            names.add(m.group(1))
    return names


def test_every_declared_symbol_is_exported(G):
    names = declared_functions()
    assert "cblas_dgemm" in names and "clapack_dgetrf" in names and "atlg_dgemm" in names
    L = G.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    # the binding declares a signature for each of them
    assert names <= set(G.exported_names()) | {"cblas_xerbla"}


def test_cblas_dgemm_argument_errors(G):
    a = np.zeros((4, 4), order="F")
    with G.XerblaTrap() as tr:
        G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, -1, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 4)
        G.cblas_dgemm(G.ColMajor, 999, G.NoTrans, 4, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 4)
        G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, 4, 4, 4, 1.0, a, 3, a, 4, 0.0, a, 4)
        G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, 4, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 2)
        G.cblas_dgemm(7, G.NoTrans, G.NoTrans, 4, 4, 4, 1.0, a, 4, a, 4, 0.0, a, 4)
        # several bad arguments: the smallest parameter number is reported
        G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, -1, -2, 4, 1.0, a, 0, a, 4, 0.0, a, 4)
    assert [c[0] for c in tr.calls] == [4, 2, 9, 14, 1, 4]
    assert all(c[1] == "cblas_dgemm" for c in tr.calls)


def test_cblas_dtrsm_dsyrk_argument_errors(G):
    a = np.zeros((4, 4), order="F")
    with G.XerblaTrap() as tr:
        G.cblas_dtrsm(G.ColMajor, 5, G.Lower, G.NoTrans, G.Unit, 4, 4, 1.0, a, 4, a, 4)
        G.cblas_dtrsm(G.ColMajor, G.Left, 5, G.NoTrans, G.Unit, 4, 4, 1.0, a, 4, a, 4)
        G.cblas_dtrsm(G.ColMajor, G.Left, G.Lower, G.NoTrans, 5, 4, 4, 1.0, a, 4, a, 4)
        G.cblas_dtrsm(G.ColMajor, G.Left, G.Lower, G.NoTrans, G.Unit, 4, -4, 1.0, a, 4, a, 4)
        G.cblas_dtrsm(G.ColMajor, G.Left, G.Lower, G.NoTrans, G.Unit, 4, 4, 1.0, a, 4, a, 3)
        G.cblas_dsyrk(G.ColMajor, G.Lower, G.NoTrans, 4, 4, 1.0, a, 3, 0.0, a, 4)
        G.cblas_dsyrk(G.ColMajor, 1, G.NoTrans, 4, 4, 1.0, a, 4, 0.0, a, 4)
    assert [c[0] for c in tr.calls] == [2, 3, 5, 7, 12, 8, 2]


def test_clapack_argument_errors_return_negative(G):
    a = np.zeros((4, 4), order="F")
    b = np.zeros((4, 1), order="F")
    ipiv = np.zeros(4, dtype=np.int32)
    with G.XerblaTrap() as tr:
        assert G.clapack_dgetrf(G.ColMajor, 4, 4, a, 3, ipiv) == -6
        assert G.clapack_dgetrf(G.ColMajor, -1, 4, a, 4, ipiv) == -2
        assert G.clapack_dgetrs(G.ColMajor, G.NoTrans, 4, 1, a, 4, ipiv, b, 2) == -9
        assert G.clapack_dgesv(G.ColMajor, 4, -1, a, 4, ipiv, b, 4) == -3
        assert G.clapack_dpotrf(G.ColMajor, 3, 4, a, 4) == -2
        assert G.clapack_dpotrs(G.ColMajor, G.Lower, 4, 1, a, 4, b, 1) == -8
        assert G.clapack_dposv(G.ColMajor, G.Lower, 4, 1, a, 1, b, 4) == -6
    assert [c[1] for c in tr.calls] == ["clapack_dgetrf", "clapack_dgetrf", "clapack_dgetrs",
                                       "clapack_dgesv", "clapack_dpotrf", "clapack_dpotrs",
                                       "clapack_dposv"]


def test_device_api_rejects_bad_arguments_without_gpu(G):
    L = G.lib()
    nul = ctypes.c_void_p(0)
    assert L.atlg_dgemm(999, G.NoTrans, 4, 4, 4, 1.0, nul, 4, nul, 4, 0.0, nul, 4, nul) == -1
    assert L.atlg_dgemm(G.NoTrans, G.NoTrans, 4, 4, 4, 1.0, nul, 3, nul, 4, 0.0, nul, 4, nul) == -8
    assert L.atlg_dtrsm(G.Left, G.Lower, G.NoTrans, G.Unit, 4, 4, 1.0, nul, 4, nul, 3, nul) == -11
    assert L.atlg_dgetrf(4, 4, nul, 4, nul, nul, nul) == -6
    assert L.atlg_dpotrf(G.Lower, 4, nul, 2, nul, nul) == -4


def test_mixed_precision_argument_errors(G):
    """Precision selectors outside their ranges are rejected on the host."""
    L = G.lib()
    c = ctypes
    s = c.c_void_p(0)
    p = c.c_void_p(16)
    it, inn = c.c_int(0), c.c_int(0)
    assert L.atlg_sgemm_bf16(4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, 3, s) == -12
    assert L.atlg_sgemm_bf16(4, 4, 4, 1.0, p, 4, p, 4, 0.0, p, 4, -1, s) == -12
    assert L.atlg_sgetrf(4, 4, p, 4, p, p, 4, s) == -7
    assert L.atlg_sgetrf(4, 4, p, 4, p, p, -1, s) == -7
    assert L.atlg_dsgesv(4, 1, p, 4, p, p, 4, p, 4, p, 30, 4, 0, c.byref(it), c.byref(inn),
                         s) == -12
    assert L.atlg_dsgesv(4, 1, p, 4, p, p, 4, p, 4, p, 30, 3, 2, c.byref(it), c.byref(inn),
                         s) == -13


def test_batched_argument_errors(G):
    """The batched entry points check every system's order and leading
    dimensions on the host before any device work."""
    L = G.lib()
    c = ctypes
    n = (c.c_int * 2)(4, -1)
    lda = (c.c_int * 2)(4, 4)
    ptrs = (c.c_void_p * 2)(16, 16)
    info = c.c_void_p(16)
    s = c.c_void_p(0)
    assert L.atlg_dpotrf_batched(-1, n, ptrs, lda, info, s) == -1
    assert L.atlg_dpotrf_batched(2, n, ptrs, lda, info, s) == -2
    n[1] = 5
    assert L.atlg_dpotrf_batched(2, n, ptrs, lda, info, s) == -4
    assert L.atlg_dpotrf_batched(2, n, ptrs, lda, None, s) == -4
    lda[1] = 5
    assert L.atlg_dpotrf_batched(2, n, ptrs, lda, None, s) == -5
    ldb = (c.c_int * 2)(4, 4)
    assert L.atlg_dposv_batched(2, n, ptrs, lda, ptrs, ldb, 4, info, s) == -6
    ldb[1] = 5
    assert L.atlg_dposv_batched(2, n, ptrs, lda, ptrs, ldb, -1, info, s) == -7
    assert L.atlg_dposv_batched(2, n, ptrs, lda, None, ldb, 4, info, s) == -5
    assert L.atlg_dsposv_batched(2, n, ptrs, lda, ptrs, ldb, ptrs, ldb, 0, 30, info, info,
                                 s) == -9
    assert L.atlg_dsposv_batched(2, n, ptrs, lda, ptrs, ldb, ptrs, ldb, 4, 30, info, None,
                                 s) == -12
    assert L.atlg_dpotrf_batched(0, n, ptrs, lda, None, s) == 0
    assert L.atlg_batched_tune_set(2, 64, 2) == -1
    assert L.atlg_batched_tune_set(0, 64, 0) == -3
    t, w = c.c_int(0), c.c_int(0)
    assert L.atlg_batched_tune_get(0, c.byref(t), c.byref(w)) == 0
    assert w.value >= 1
