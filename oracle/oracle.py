"""ctypes wrapper for the CPU oracle (TEST INFRASTRUCTURE ONLY).

The oracle restates ATLAS 3.11.42's tester generator, norms and residual
formulas plus textbook reference routines (see atl_oracle.c for citations).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker.  All matrices are column-major numpy
float64 arrays: an (lda, n) array is handled as Fortran-ordered storage by
passing arrays of shape (n, lda) in C order and treating rows as columns.
Helpers below hide that: they take/return arrays of shape (m, n) in
Fortran order (order='F').
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i = ctypes.c_int
_d = ctypes.c_double


def build():
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        sig = {
            "orc_srand": (None, [_i]),
            "orc_rand": (_i, []),
            "orc_dumb_rand": (_d, []),
            "orc_rand_words": (None, [_i, _i, _ip]),
            "orc_gegen": (None, [_i, _i, _dp, _i, _i]),
            "orc_gechkgap": (_i, [_i, _i, _dp, _i]),
            "orc_trddom": (None, [_i, _i, _dp, _i]),
            "orc_lltgen": (None, [_i, _i, _dp, _i, _i]),
            "orc_epsilon": (_d, []),
            "orc_genrm1": (_d, [_i, _i, _dp, _i]),
            "orc_gediffnrm1": (_d, [_i, _i, _dp, _i, _dp, _i]),
            "orc_trnrm1": (_d, [_i, _i, _i, _dp, _i]),
            "orc_lltnrm1": (_d, [_i, _i, _dp, _i]),
            "orc_infnrm": (_d, [_i, _dp, _i]),
            "orc_refgemm": (None, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
            "orc_refsyrk": (None, [_i, _i, _i, _i, _d, _dp, _i, _d, _dp, _i]),
            "orc_reftrsm": (None, [_i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]),
            "orc_idamax": (_i, [_i, _dp, _i]),
            "orc_getf2": (_i, [_i, _i, _dp, _i, _ip]),
            "orc_laswp": (None, [_i, _dp, _i, _i, _i, _ip, _i]),
            "orc_potf2": (_i, [_i, _i, _dp, _i]),
            "orc_getrs": (None, [_i, _i, _i, _dp, _i, _ip, _dp, _i]),
            "orc_potrs": (None, [_i, _i, _i, _dp, _i, _dp, _i]),
            "orc_gemm_resid": (_d, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d,
                                    _dp, _i, _dp, _i, _dp, _i]),
            "orc_gemm_elem_check": (_i, [_i, _i, _i, _d, _d, _dp, _i, _dp, _i, _dp]),
            "orc_trsm_resid": (_d, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _dp,
                                    _i, _dp, _i]),
            "orc_lu_resid": (_d, [_i, _i, _dp, _i, _dp, _i, _ip]),
            "orc_llt_resid": (_d, [_i, _i, _dp, _i, _dp, _i]),
            "orc_slv_resid": (_d, [_i, _i, _d, _dp, _i, _dp, _i]),
            "orc_npvts": (_i, [_i, _ip]),
            "orc_time_refgemm_slab": (_d, [_i, _i, _i, _dp, _i, _dp, _i, _dp, _i]),
            "orc_time_refgemm_slab_mt": (_d, [_i, _i, _i, _dp, _i, _dp, _i, _dp, _i, _i]),
            "orc_reftrmm": (None, [_i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]),
            "orc_refsymm": (None, [_i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
            "orc_refsyr2k": (None, [_i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
            "orc_reftrtri": (_i, [_i, _i, _i, _dp, _i]),
            "orc_reflauum": (None, [_i, _i, _dp, _i]),
            "orc_blat_dbeg": (_d, [_ip]),
            "orc_blat_dmake": (None, [_i, _i, _i, _i, _i, _dp, _i, _dp, _i, _ip, _d]),
            "orc_blat_dmmch": (_i, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i, _dp,
                                    _i, _d, ctypes.POINTER(ctypes.c_double),
                                    ctypes.POINTER(ctypes.c_double)]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB


def _p(a):
    """double* into a Fortran-ordered (or 1-D) float64 array."""
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim == 1)
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_ip)


def _ld(a):
    return max(1, a.shape[0]) if a.ndim == 2 else max(1, a.shape[0])


def fmat(m, n):
    return np.zeros((m, n), dtype=np.float64, order="F")


# ---- generator ---------------------------------------------------------

def rand_words(seed, n):
    out = np.zeros(n, dtype=np.int32)
    lib().orc_rand_words(seed, n, _pi(out))
    return out


def gegen(m, n, seed, lda=None):
    """ATL_gegen into an (lda, n) Fortran array; returns the full array."""
    lda = m if lda is None else lda
    a = fmat(max(lda, 1), n)
    if m > 0 and n > 0:
        lib().orc_gegen(m, n, _p(a), max(lda, 1), seed)
    return a


def gechkgap(a, m):
    return lib().orc_gechkgap(m, a.shape[1], _p(a), a.shape[0])


def trddom(a, uplo, n):
    lib().orc_trddom(uplo, n, _p(a), a.shape[0])


def lltgen(uplo, n, seed, lda=None):
    lda = n if lda is None else lda
    a = fmat(max(lda, 1), n)
    if n > 0:
        lib().orc_lltgen(uplo, n, _p(a), max(lda, 1), seed)
    return a


# ---- scalars / norms ------------------------------------------------------

def epsilon():
    return lib().orc_epsilon()


def genrm1(a, m=None, n=None):
    m = a.shape[0] if m is None else m
    n = a.shape[1] if n is None else n
    return lib().orc_genrm1(m, n, _p(a), a.shape[0])


def trnrm1(uplo, diag, a, n):
    return lib().orc_trnrm1(uplo, diag, n, _p(a), a.shape[0])


def lltnrm1(uplo, a, n):
    return lib().orc_lltnrm1(uplo, n, _p(a), a.shape[0])


# ---- textbook routines (in place on Fortran arrays) -----------------------

def refgemm(ta, tb, m, n, k, alpha, a, b, beta, c):
    lib().orc_refgemm(ta, tb, m, n, k, alpha, _p(a), _ld(a), _p(b), _ld(b), beta, _p(c), _ld(c))


def refsyrk(uplo, trans, n, k, alpha, a, beta, c):
    lib().orc_refsyrk(uplo, trans, n, k, alpha, _p(a), _ld(a), beta, _p(c), _ld(c))


def reftrsm(side, uplo, ta, diag, m, n, alpha, a, b):
    lib().orc_reftrsm(side, uplo, ta, diag, m, n, alpha, _p(a), _ld(a), _p(b), _ld(b))


def reftrmm(side, uplo, ta, diag, m, n, alpha, a, b):
    lib().orc_reftrmm(side, uplo, ta, diag, m, n, alpha, _p(a), _ld(a), _p(b), _ld(b))


def refsymm(side, uplo, m, n, alpha, a, b, beta, c):
    lib().orc_refsymm(side, uplo, m, n, alpha, _p(a), _ld(a), _p(b), _ld(b), beta, _p(c), _ld(c))


def refsyr2k(uplo, trans, n, k, alpha, a, b, beta, c):
    lib().orc_refsyr2k(uplo, trans, n, k, alpha, _p(a), _ld(a), _p(b), _ld(b), beta, _p(c),
                       _ld(c))


def reftrtri(uplo, diag, a, n=None):
    n = a.shape[1] if n is None else n
    return lib().orc_reftrtri(uplo, diag, n, _p(a), _ld(a))


def reflauum(uplo, a, n=None):
    n = a.shape[1] if n is None else n
    lib().orc_reflauum(uplo, n, _p(a), _ld(a))


def idamax(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_idamax(x.size, _p(x), 1)


def getf2(a, m=None, n=None):
    m = a.shape[0] if m is None else m
    n = a.shape[1] if n is None else n
    ipiv = np.zeros(max(min(m, n), 1), dtype=np.int32)
    info = lib().orc_getf2(m, n, _p(a), _ld(a), _pi(ipiv))
    return ipiv[:min(m, n)], info


def laswp(a, k1, k2, ipiv, inci, n=None):
    n = a.shape[1] if n is None else n
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_laswp(n, _p(a), _ld(a), k1, k2, _pi(ipiv), inci)


def potf2(uplo, a, n=None):
    n = a.shape[1] if n is None else n
    return lib().orc_potf2(uplo, n, _p(a), _ld(a))


def getrs(trans, a, ipiv, b, n=None, nrhs=None):
    n = a.shape[1] if n is None else n
    nrhs = b.shape[1] if nrhs is None else nrhs
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    lib().orc_getrs(trans, n, nrhs, _p(a), _ld(a), _pi(ipiv), _p(b), _ld(b))


def potrs(uplo, a, b, n=None, nrhs=None):
    n = a.shape[1] if n is None else n
    nrhs = b.shape[1] if nrhs is None else nrhs
    lib().orc_potrs(uplo, n, nrhs, _p(a), _ld(a), _p(b), _ld(b))


# ---- tester residuals -----------------------------------------------------

def gemm_resid(ta, tb, m, n, k, alpha, a, b, beta, c0, ctest, ctrust):
    return lib().orc_gemm_resid(ta, tb, m, n, k, alpha, _p(a), _ld(a), _p(b), _ld(b), beta,
                                _p(c0), _ld(c0), _p(ctest), _ld(ctest), _p(ctrust), _ld(ctrust))


def gemm_elem_check(m, n, k, alpha, beta, ctest, ctrust):
    mx = ctypes.c_double(0.0)
    nbad = lib().orc_gemm_elem_check(m, n, k, alpha, beta, _p(ctest), _ld(ctest), _p(ctrust),
                                     _ld(ctrust), ctypes.byref(mx))
    return nbad, mx.value


def trsm_resid(side, uplo, diag, m, n, alpha, a, b0, btest, btrust):
    return lib().orc_trsm_resid(side, uplo, diag, m, n, alpha, _p(a), _ld(a), _p(b0), _ld(b0),
                                _p(btest), _ld(btest), _p(btrust), _ld(btrust))


def lu_resid(m, n, a0, lu, ipiv):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return lib().orc_lu_resid(m, n, _p(a0), _ld(a0), _p(lu), _ld(lu), _pi(ipiv))


def llt_resid(uplo, n, a0, l):
    return lib().orc_llt_resid(uplo, n, _p(a0), _ld(a0), _p(l), _ld(l))


def slv_resid(n, nrhs, nrmA, xc, x):
    return lib().orc_slv_resid(n, nrhs, nrmA, _p(xc), _ld(xc), _p(x), _ld(x))


def npvts(ipiv):
    ipiv = np.ascontiguousarray(ipiv, dtype=np.int32)
    return lib().orc_npvts(ipiv.size, _pi(ipiv))


def time_refgemm_slab(m, ncols, k, a, b, c):
    return lib().orc_time_refgemm_slab(m, ncols, k, _p(a), _ld(a), _p(b), _ld(b), _p(c), _ld(c))


def time_refgemm_slab_mt(m, ncols, k, a, b, c, nthreads):
    return lib().orc_time_refgemm_slab_mt(m, ncols, k, _p(a), _ld(a), _p(b), _ld(b), _p(c), _ld(c),
                                          nthreads)


# ---- netlib CBLAS level-3 tester (c_dblat3.f) machinery --------------------

class BlatGen:
    """DBEG's state; reset=True restarts the sequence (RESET in the Fortran)."""

    def __init__(self):
        self.st = np.array([0, 0, 0, 1], dtype=np.int32)

    def reset(self):
        self.st[3] = 1

    def dbeg(self):
        return lib().orc_blat_dbeg(_pi(self.st))


BLAT_GE, BLAT_SY, BLAT_TR = 0, 1, 2


def blat_dmake(gen, typ, uplo, diag, m, n, nmax, lda, transl=0.0):
    """DMAKE: returns (A as an nmax-leading Fortran array, AA with leading
    dimension lda)."""
    a = fmat(max(nmax, 1), max(n, 1))
    aa = fmat(max(lda, 1), max(n, 1))
    lib().orc_blat_dmake(typ, uplo, diag, m, n, _p(a), max(nmax, 1), _p(aa), max(lda, 1),
                         _pi(gen.st), transl)
    return a, aa


def blat_eps():
    """c_dblat3's EPS: twice LAPACK's dlamch('E') (2^-53), i.e. 2^-52."""
    return 2.0 * 2.0 ** -53


def blat_dmmch(ta, tb, m, n, kk, alpha, a, b, beta, c, cc, eps=None):
    """DMMCH: (fatal, err of the last column, largest column err)."""
    eps = blat_eps() if eps is None else eps
    err, emax = ctypes.c_double(0.0), ctypes.c_double(0.0)
    fatal = lib().orc_blat_dmmch(ta, tb, m, n, kk, alpha, _p(a), _ld(a), _p(b), _ld(b), beta,
                                 _p(c), _ld(c), _p(cc), _ld(cc), eps, ctypes.byref(err),
                                 ctypes.byref(emax))
    return bool(fatal), err.value, emax.value
