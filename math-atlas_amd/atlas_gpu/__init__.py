"""Python binding of libatlas_gpu.so (the MI355X FP64 level-3 library).

Two layers, both straight ctypes calls into the C ABI declared in
include/*.h:

* the drop-in host-pointer entry points (cblas_dgemm, clapack_dgetrf, ...),
  which take numpy arrays and behave like ATLAS's (synchronous, in place);
* the device-pointer atlg_* API, which takes torch CUDA tensors holding
  column-major matrices and queues work on a HIP stream.

There is no fallback: if the shared library is missing, loading raises.
Column-major matrices in torch are stored as a (cols, ld) row-major tensor,
so element (i, j) is t[j, i]; `colmajor` and `from_fortran` build them.
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.dirname(_HERE)
LIB_PATH = os.path.join(PKG_DIR, "libatlas_gpu.so")

RowMajor, ColMajor = 101, 102
NoTrans, Trans, ConjTrans = 111, 112, 113
Upper, Lower = 121, 122
NonUnit, Unit = 131, 132
Left, Right = 141, 142
EHIP = -1000
EWAIT = -1001
ESIZE = -1002
FAIL_PANEL, FAIL_TRSV = 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_i, _l, _d, _v = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
_ul = ctypes.c_ulong
_cp = ctypes.c_char_p

XERBLA_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p)
FAILURE_HOOK = ctypes.CFUNCTYPE(None, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p)

_SIGS = {
    # drop-in host API
    "cblas_dgemm": (None, [_i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
    "cblas_dtrsm": (None, [_i, _i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]),
    "cblas_dsyrk": (None, [_i, _i, _i, _i, _i, _d, _dp, _i, _d, _dp, _i]),
    "clapack_dgetrf": (_i, [_i, _i, _i, _dp, _i, _ip]),
    "clapack_dgetrs": (_i, [_i, _i, _i, _i, _dp, _i, _ip, _dp, _i]),
    "clapack_dgesv": (_i, [_i, _i, _i, _dp, _i, _ip, _dp, _i]),
    "clapack_dpotrf": (_i, [_i, _i, _i, _dp, _i]),
    "clapack_dpotrs": (_i, [_i, _i, _i, _i, _dp, _i, _dp, _i]),
    "clapack_dposv": (_i, [_i, _i, _i, _i, _dp, _i, _dp, _i]),
    "ATL_dgemm": (None, [_i, _i, _ul, _ul, _ul, _d, _dp, _ul, _dp, _ul, _d, _dp, _ul]),
    "ATL_dsyrk": (None, [_i, _i, _ul, _ul, _d, _dp, _ul, _d, _dp, _ul]),
    "ATL_dtrsm": (None, [_i, _i, _i, _i, _ul, _ul, _d, _dp, _ul, _dp, _ul]),
    "ATL_dgetrf": (_i, [_i, _i, _i, _dp, _i, _ip]),
    "ATL_dgetrfC": (_i, [_i, _i, _dp, _i, _ip]),
    "ATL_dgetrfR": (_i, [_i, _i, _dp, _i, _ip]),
    "ATL_dgetrs": (None, [_i, _i, _i, _i, _dp, _i, _ip, _dp, _i]),
    "ATL_dlaswp": (None, [_i, _dp, _i, _i, _i, _ip, _i]),
    "ATL_dpotrf": (_i, [_i, _i, _i, _dp, _i]),
    "ATL_dpotrfL": (_i, [_i, _dp, _i]),
    "ATL_dpotrfU": (_i, [_i, _dp, _i]),
    "ATL_dpotrs": (None, [_i, _i, _i, _i, _dp, _i, _dp, _i]),
    "cblas_dtrmm": (None, [_i, _i, _i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i]),
    "cblas_dsymm": (None, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
    "cblas_dsyr2k": (None, [_i, _i, _i, _i, _i, _d, _dp, _i, _dp, _i, _d, _dp, _i]),
    "clapack_dtrtri": (_i, [_i, _i, _i, _i, _dp, _i]),
    "clapack_dpotri": (_i, [_i, _i, _i, _dp, _i]),
    "clapack_dlauum": (_i, [_i, _i, _i, _dp, _i]),
    "ATL_dtrmm": (None, [_i, _i, _i, _i, _ul, _ul, _d, _dp, _ul, _dp, _ul]),
    "ATL_dsymm": (None, [_i, _i, _ul, _ul, _d, _dp, _ul, _dp, _ul, _d, _dp, _ul]),
    "ATL_dsyr2k": (None, [_i, _i, _ul, _ul, _d, _dp, _ul, _dp, _ul, _d, _dp, _ul]),
    "ATL_dtrtri": (_i, [_i, _i, _i, _i, _dp, _i]),
    "ATL_dlauum": (None, [_i, _i, _i, _dp, _i]),
    "ATL_dpotri": (_i, [_i, _i, _i, _dp, _i]),
    # Fortran 77 interface (every argument by reference)
    "xerbla_": (None, [ctypes.c_char_p, _ip, ctypes.c_size_t]),
    "dgemm_": (None, [_cp, _cp, _ip, _ip, _ip, _dp, _dp, _ip, _dp, _ip, _dp, _dp, _ip]),
    "dsyrk_": (None, [_cp, _cp, _ip, _ip, _dp, _dp, _ip, _dp, _dp, _ip]),
    "dtrsm_": (None, [_cp, _cp, _cp, _cp, _ip, _ip, _dp, _dp, _ip, _dp, _ip]),
    "dtrmm_": (None, [_cp, _cp, _cp, _cp, _ip, _ip, _dp, _dp, _ip, _dp, _ip]),
    "dsymm_": (None, [_cp, _cp, _ip, _ip, _dp, _dp, _ip, _dp, _ip, _dp, _dp, _ip]),
    "dsyr2k_": (None, [_cp, _cp, _ip, _ip, _dp, _dp, _ip, _dp, _ip, _dp, _dp, _ip]),
    "dgetrf_": (None, [_ip, _ip, _dp, _ip, _ip, _ip]),
    "dgetrs_": (None, [_cp, _ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]),
    "dgesv_": (None, [_ip, _ip, _dp, _ip, _ip, _dp, _ip, _ip]),
    "dpotrf_": (None, [_cp, _ip, _dp, _ip, _ip]),
    "dpotrs_": (None, [_cp, _ip, _ip, _dp, _ip, _dp, _ip, _ip]),
    "dposv_": (None, [_cp, _ip, _ip, _dp, _ip, _dp, _ip, _ip]),
    "dtrtri_": (None, [_cp, _cp, _ip, _dp, _ip, _ip]),
    "dpotri_": (None, [_cp, _ip, _dp, _ip, _ip]),
    "dlauum_": (None, [_cp, _ip, _dp, _ip, _ip]),
    "dlaswp_": (None, [_ip, _dp, _ip, _ip, _ip, _ip, _ip]),
    # device API
    "atlg_init": (_i, []),
    "atlg_last_error": (ctypes.c_char_p, []),
    "atlg_set_xerbla_hook": (None, [XERBLA_HOOK]),
    "atlg_stream_status": (_i, [_v, _ip]),
    "atlg_set_failure_hook": (None, [FAILURE_HOOK]),
    "atlg_set_wait_limit": (None, [_l]),
    "atlg_dgemm": (_i, [_i, _i, _l, _l, _l, _d, _v, _l, _v, _l, _d, _v, _l, _v]),
    "atlg_dsyrk": (_i, [_i, _i, _l, _l, _d, _v, _l, _d, _v, _l, _v]),
    "atlg_dtrsm": (_i, [_i, _i, _i, _i, _l, _l, _d, _v, _l, _v, _l, _v]),
    "atlg_dtrsm_noalloc": (_i, [_i, _i, _i, _i, _l, _l, _d, _v, _l, _v, _l, _v]),
    "atlg_dlaswp": (_i, [_l, _v, _l, _i, _i, _v, _i, _v]),
    "atlg_dgetrf": (_i, [_l, _l, _v, _l, _v, _v, _v]),
    "atlg_dgetrs": (_i, [_i, _l, _l, _v, _l, _v, _v, _l, _v]),
    "atlg_dpotrf": (_i, [_i, _l, _v, _l, _v, _v]),
    "atlg_dpotrs": (_i, [_i, _l, _l, _v, _l, _v, _l, _v]),
    "atlg_dtrmm": (_i, [_i, _i, _i, _i, _l, _l, _d, _v, _l, _v, _l, _v]),
    "atlg_dsymm": (_i, [_i, _i, _l, _l, _d, _v, _l, _v, _l, _d, _v, _l, _v]),
    "atlg_dsyr2k": (_i, [_i, _i, _l, _l, _d, _v, _l, _v, _l, _d, _v, _l, _v]),
    "atlg_dtrtri": (_i, [_i, _i, _l, _v, _l, _v, _v]),
    "atlg_dlauum": (_i, [_i, _l, _v, _l, _v]),
    "atlg_dpotri": (_i, [_i, _l, _v, _l, _v, _v]),
    "atlg_dpotrf_adjoint": (_i, [_l, _v, _l, _v, _l, _v]),
    "atlg_dtrsm_adjoint": (_i, [_l, _l, _v, _l, _v, _l, _v, _l, _v, _l, _v]),
    "atlg_dlogdet_chol": (_i, [_l, _v, _l, _v, _v]),
    "atlg_dlogdet_chol_adjoint": (_i, [_l, _v, _l, _v, _v, _l, _v]),
    "atlg_sgemm": (_i, [_i, _i, _l, _l, _l, ctypes.c_float, _v, _l, _v, _l, ctypes.c_float, _v,
                        _l, _v]),
    "atlg_sgemm_bf16": (_i, [_l, _l, _l, ctypes.c_float, _v, _l, _v, _l, ctypes.c_float, _v, _l,
                             _i, _v]),
    "atlg_sgetrf": (_i, [_l, _l, _v, _l, _v, _v, _i, _v]),
    "atlg_sgetrs": (_i, [_i, _l, _l, _v, _l, _v, _v, _l, _v]),
    "atlg_dsgesv": (_i, [_l, _l, _v, _l, _v, _v, _l, _v, _l, _v, _i, _i, _i, _ip, _ip, _v]),
    "atlg_dpotrf_batched": (_i, [_i, _ip, _v, _ip, _v, _v]),
    "atlg_dpotrs_batched": (_i, [_i, _ip, _v, _ip, _v, _ip, _i, _v]),
    "atlg_dposv_batched": (_i, [_i, _ip, _v, _ip, _v, _ip, _i, _v, _v]),
    "atlg_spotrf_batched": (_i, [_i, _ip, _v, _ip, _v, _v]),
    "atlg_sposv_batched": (_i, [_i, _ip, _v, _ip, _v, _ip, _i, _v, _v]),
    "atlg_dsposv_batched": (_i, [_i, _ip, _v, _ip, _v, _ip, _v, _ip, _i, _i, _v, _v, _v]),
    "atlg_batched_tune_get": (_i, [_i, _ip, _ip]),
    "atlg_batched_tune_set": (_i, [_i, _i, _i]),
}

_LIB = None


def lib():
    """Loads libatlas_gpu.so (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                "libatlas_gpu.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        # bounded-wait failures of the host entry points raise AtlgWaitError
        # (on the calling thread) instead of the C default, print and abort
        L.atlg_set_failure_hook(_FAILURE_CB)
        _LIB = L
    return _LIB


def exported_names():
    return list(_SIGS)


class AtlgError(RuntimeError):
    pass


class AtlgWaitError(AtlgError):
    """A bounded wait inside a kernel gave up (ATLG_FAIL_* bits in .bits);
    the results of the call are unspecified."""

    def __init__(self, msg, bits):
        super().__init__(msg)
        self.bits = bits


_failures = threading.local()


def _on_failure(rout, bits, msg):
    _failures.last = (rout.decode(), bits, msg.decode())


_FAILURE_CB = FAILURE_HOOK(_on_failure)


def _host(fn, *args):
    """Calls a host-pointer entry point; raises AtlgWaitError when its
    device work reported a bounded-wait failure."""
    _failures.last = None
    r = fn(*args)
    f = getattr(_failures, "last", None)
    if f is not None:
        _failures.last = None
        raise AtlgWaitError("%s: %s" % (f[0], f[2]), f[1])
    return r


def set_wait_limit(polls):
    """Polls before a bounded wait gives up (<= 0: the defaults)."""
    lib().atlg_set_wait_limit(int(polls))


def stream_status(stream=None):
    """Synchronizes the stream and returns (and clears) its failure bits."""
    st = ctypes.c_int(0)
    _check(lib().atlg_stream_status(_stream(stream), ctypes.byref(st)), "atlg_stream_status")
    return st.value


def check_stream(stream=None):
    """Raises AtlgWaitError if work queued on the stream reported a
    bounded-wait failure (synchronizes the stream)."""
    bits = stream_status(stream)
    if bits:
        raise AtlgWaitError("a bounded wait gave up on this stream (bits %d)" % bits, bits)


def _check(rc, what):
    if rc == EHIP:
        raise AtlgError("%s: HIP failure: %s" % (what, lib().atlg_last_error().decode()))
    if rc == EWAIT:
        raise AtlgWaitError("%s: %s" % (what, lib().atlg_last_error().decode()), 0)
    if rc != 0:
        raise AtlgError("%s: illegal argument %d" % (what, -rc))


# ---------------------------------------------------------------- xerbla hook

class XerblaTrap:
    """Context manager: records cblas_xerbla calls instead of exiting."""

    def __init__(self):
        self.calls = []
        self._cb = XERBLA_HOOK(self._hook)

    def _hook(self, p, rout, msg):
        self.calls.append((p, rout.decode(), msg.decode()))

    def __enter__(self):
        lib().atlg_set_xerbla_hook(self._cb)
        return self

    def __exit__(self, *exc):
        lib().atlg_set_xerbla_hook(XERBLA_HOOK())
        return False


# ---------------------------------------------------------------- host API

def _p(a):
    if a is None:
        return ctypes.cast(None, _dp)
    assert a.dtype == np.float64
    return a.ctypes.data_as(_dp)


def _pi(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def cblas_dgemm(order, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _host(lib().cblas_dgemm, order, ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def cblas_dtrsm(order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    _host(lib().cblas_dtrsm, order, side, uplo, ta, diag, m, n, alpha, _p(a), lda, _p(b), ldb)


def cblas_dsyrk(order, uplo, trans, n, k, alpha, a, lda, beta, c, ldc):
    _host(lib().cblas_dsyrk, order, uplo, trans, n, k, alpha, _p(a), lda, beta, _p(c), ldc)


def clapack_dgetrf(order, m, n, a, lda, ipiv):
    return _host(lib().clapack_dgetrf, order, m, n, _p(a), lda, _pi(ipiv))


def clapack_dgetrs(order, trans, n, nrhs, a, lda, ipiv, b, ldb):
    return _host(lib().clapack_dgetrs, order, trans, n, nrhs, _p(a), lda, _pi(ipiv), _p(b), ldb)


def clapack_dgesv(order, n, nrhs, a, lda, ipiv, b, ldb):
    return _host(lib().clapack_dgesv, order, n, nrhs, _p(a), lda, _pi(ipiv), _p(b), ldb)


def clapack_dpotrf(order, uplo, n, a, lda):
    return _host(lib().clapack_dpotrf, order, uplo, n, _p(a), lda)


def clapack_dpotrs(order, uplo, n, nrhs, a, lda, b, ldb):
    return _host(lib().clapack_dpotrs, order, uplo, n, nrhs, _p(a), lda, _p(b), ldb)


def clapack_dposv(order, uplo, n, nrhs, a, lda, b, ldb):
    return _host(lib().clapack_dposv, order, uplo, n, nrhs, _p(a), lda, _p(b), ldb)


def ATL_dlaswp(n, a, lda, k1, k2, ipiv, inci):
    _host(lib().ATL_dlaswp, n, _p(a), lda, k1, k2, _pi(ipiv), inci)


def cblas_dtrmm(order, side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    _host(lib().cblas_dtrmm, order, side, uplo, ta, diag, m, n, alpha, _p(a), lda, _p(b), ldb)


def cblas_dsymm(order, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc):
    _host(lib().cblas_dsymm, order, side, uplo, m, n, alpha, _p(a), lda, _p(b), ldb, beta, _p(c),
          ldc)


def cblas_dsyr2k(order, uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _host(lib().cblas_dsyr2k, order, uplo, trans, n, k, alpha, _p(a), lda, _p(b), ldb, beta,
          _p(c), ldc)


def clapack_dtrtri(order, uplo, diag, n, a, lda):
    return _host(lib().clapack_dtrtri, order, uplo, diag, n, _p(a), lda)


def clapack_dpotri(order, uplo, n, a, lda):
    return _host(lib().clapack_dpotri, order, uplo, n, _p(a), lda)


def clapack_dlauum(order, uplo, n, a, lda):
    return _host(lib().clapack_dlauum, order, uplo, n, _p(a), lda)


class F77:
    """The Fortran 77 entry points (dgemm_, dgetrf_, ...) called the way a
    Fortran program calls them: every scalar by reference, characters as
    one-byte strings, 1-based pivots."""

    @staticmethod
    def _i(v):
        return ctypes.byref(ctypes.c_int(int(v)))

    @staticmethod
    def _d(v):
        return ctypes.byref(ctypes.c_double(float(v)))

    @staticmethod
    def _c(v):
        return v.encode() if isinstance(v, str) else v

    @classmethod
    def call(cls, name, *args):
        """args: ('c', 'N'), ('i', 5), ('d', 1.5), ('a', ndarray) or
        ('info', None) (an int returned)."""
        out, conv = None, []
        for kind, v in args:
            if kind == "c":
                conv.append(cls._c(v))
            elif kind == "i":
                conv.append(cls._i(v))
            elif kind == "d":
                conv.append(cls._d(v))
            elif kind == "a":
                conv.append(_p(v) if v.dtype == np.float64 else _pi(v))
            elif kind == "info":
                out = ctypes.c_int(0)
                conv.append(ctypes.byref(out))
        _host(getattr(lib(), name), *conv)
        return None if out is None else out.value


def ATL_dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    _host(lib().ATL_dgemm, ta, tb, m, n, k, alpha, _p(a), lda, _p(b), ldb, beta, _p(c), ldc)


def ATL_dtrsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb):
    _host(lib().ATL_dtrsm, side, uplo, ta, diag, m, n, alpha, _p(a), lda, _p(b), ldb)


def ATL_dgetrf(order, m, n, a, lda, ipiv):
    return _host(lib().ATL_dgetrf, order, m, n, _p(a), lda, _pi(ipiv))


def ATL_dpotrf(order, uplo, n, a, lda):
    return _host(lib().ATL_dpotrf, order, uplo, n, _p(a), lda)


# ---------------------------------------------------------------- device API

def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("atlas_gpu device API needs a GPU (torch.cuda unavailable)")
    return torch


def colmajor(rows, cols, ld=None, device="cuda", fill=None):
    """A column-major rows x cols FP64 matrix with leading dimension ld,
    stored as a (cols, ld) torch tensor."""
    torch = _torch()
    ld = max(1, rows if ld is None else ld)
    t = torch.empty((max(cols, 1), ld), dtype=torch.float64, device=device)
    if fill is not None:
        t.fill_(fill)
    return t


def from_fortran(a, device="cuda"):
    """numpy (ld, cols) Fortran-ordered float64 -> device (cols, ld) tensor."""
    torch = _torch()
    a = np.asfortranarray(a, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(a.T)).to(device)


def to_fortran(t):
    """device (cols, ld) tensor -> numpy (ld, cols) Fortran-ordered array."""
    return np.asfortranarray(t.detach().cpu().numpy().T)


def _ptr(t):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _stream(stream):
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    _check(lib().atlg_dgemm(ta, tb, m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c),
                            ldc, _stream(stream)), "atlg_dgemm")


def dsyrk(uplo, trans, n, k, alpha, a, lda, beta, c, ldc, stream=None):
    _check(lib().atlg_dsyrk(uplo, trans, n, k, alpha, _ptr(a), lda, beta, _ptr(c), ldc,
                            _stream(stream)), "atlg_dsyrk")


def dtrsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, stream=None, check=False):
    """check=True: synchronize and raise AtlgWaitError on a bounded-wait
    failure (the few-RHS solve)."""
    _check(lib().atlg_dtrsm(side, uplo, ta, diag, m, n, alpha, _ptr(a), lda, _ptr(b), ldb,
                            _stream(stream)), "atlg_dtrsm")
    if check:
        check_stream(stream)


def dtrsm_noalloc(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, stream=None):
    """atlg_dtrsm_noalloc: dtrsm with no device allocation (order of op(A)
    <= 8192; its static workspace is shared by every call on the GPU)."""
    rc = lib().atlg_dtrsm_noalloc(side, uplo, ta, diag, m, n, alpha, _ptr(a), lda, _ptr(b), ldb,
                                  _stream(stream))
    if rc == ESIZE:
        raise AtlgError("atlg_dtrsm_noalloc: %s" % lib().atlg_last_error().decode())
    _check(rc, "atlg_dtrsm_noalloc")


def dlaswp(n, a, lda, k1, k2, ipiv, inci, stream=None):
    _check(lib().atlg_dlaswp(n, _ptr(a), lda, k1, k2, _ptr(ipiv), inci, _stream(stream)),
           "atlg_dlaswp")


def dgetrf(m, n, a, lda, ipiv, info, stream=None, check=False):
    """check=True: synchronize and raise AtlgWaitError on a bounded-wait
    failure of a panel (info then holds INT_MIN)."""
    _check(lib().atlg_dgetrf(m, n, _ptr(a), lda, _ptr(ipiv), _ptr(info), _stream(stream)),
           "atlg_dgetrf")
    if check:
        check_stream(stream)


def dgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb, stream=None, check=False):
    _check(lib().atlg_dgetrs(trans, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb,
                             _stream(stream)), "atlg_dgetrs")
    if check:
        check_stream(stream)


def dpotrf(uplo, n, a, lda, info, stream=None):
    _check(lib().atlg_dpotrf(uplo, n, _ptr(a), lda, _ptr(info), _stream(stream)), "atlg_dpotrf")


def dpotrs(uplo, n, nrhs, a, lda, b, ldb, stream=None):
    _check(lib().atlg_dpotrs(uplo, n, nrhs, _ptr(a), lda, _ptr(b), ldb, _stream(stream)),
           "atlg_dpotrs")


def dtrmm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb, stream=None):
    _check(lib().atlg_dtrmm(side, uplo, ta, diag, m, n, alpha, _ptr(a), lda, _ptr(b), ldb,
                            _stream(stream)), "atlg_dtrmm")


def dsymm(side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    _check(lib().atlg_dsymm(side, uplo, m, n, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c),
                            ldc, _stream(stream)), "atlg_dsymm")


def dsyr2k(uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    _check(lib().atlg_dsyr2k(uplo, trans, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c),
                             ldc, _stream(stream)), "atlg_dsyr2k")


def dtrtri(uplo, diag, n, a, lda, info, stream=None):
    _check(lib().atlg_dtrtri(uplo, diag, n, _ptr(a), lda, _ptr(info), _stream(stream)),
           "atlg_dtrtri")


def dlauum(uplo, n, a, lda, stream=None):
    _check(lib().atlg_dlauum(uplo, n, _ptr(a), lda, _stream(stream)), "atlg_dlauum")


def dpotri(uplo, n, a, lda, info, stream=None):
    _check(lib().atlg_dpotri(uplo, n, _ptr(a), lda, _ptr(info), _stream(stream)), "atlg_dpotri")


def dpotrf_adjoint(n, L, ldl, G, ldg, stream=None):
    _check(lib().atlg_dpotrf_adjoint(n, _ptr(L), ldl, _ptr(G), ldg, _stream(stream)),
           "atlg_dpotrf_adjoint")


def dtrsm_adjoint(n, nrhs, L, ldl, X, ldx, Xbar, ldxb, Lbar, ldlb, stream=None):
    _check(lib().atlg_dtrsm_adjoint(n, nrhs, _ptr(L), ldl, _ptr(X), ldx, _ptr(Xbar), ldxb,
                                    _ptr(Lbar), ldlb, _stream(stream)), "atlg_dtrsm_adjoint")


def dlogdet_chol(n, L, ldl, out, stream=None):
    _check(lib().atlg_dlogdet_chol(n, _ptr(L), ldl, _ptr(out), _stream(stream)),
           "atlg_dlogdet_chol")


def dlogdet_chol_adjoint(n, L, ldl, dbar, Lbar, ldlb, stream=None):
    _check(lib().atlg_dlogdet_chol_adjoint(n, _ptr(L), ldl, _ptr(dbar), _ptr(Lbar), ldlb,
                                           _stream(stream)), "atlg_dlogdet_chol_adjoint")


# ---------------------------------------------------------------- batched SPD solvers
# n, lda, ldb, ldx: sequences of ints (host); A, B, X: sequences of device
# addresses (ints, e.g. tensor.data_ptr() + byte offset) or tensors; info and
# iter: int32 device tensors of len(n).

def _ints(seq):
    a = np.ascontiguousarray(np.asarray(seq, dtype=np.int32))
    return a, a.ctypes.data_as(_ip)


def _ptrs(seq):
    if seq is None:
        return None, ctypes.c_void_p(0)
    vals = [(x.data_ptr() if hasattr(x, "data_ptr") else int(x)) for x in seq]
    a = (ctypes.c_void_p * max(1, len(vals)))(*vals)
    return a, ctypes.cast(a, ctypes.c_void_p)


def dpotrf_batched(n, A, lda, info, stream=None):
    (kn, pn), (ka, pa), (kl, pl) = _ints(n), _ptrs(A), _ints(lda)
    _check(lib().atlg_dpotrf_batched(len(kn), pn, pa, pl, _ptr(info), _stream(stream)),
           "atlg_dpotrf_batched")


def dpotrs_batched(n, A, lda, B, ldb, nrhs, stream=None):
    (kn, pn), (ka, pa), (kl, pl), (kb, pb), (klb, plb) = (_ints(n), _ptrs(A), _ints(lda),
                                                          _ptrs(B), _ints(ldb))
    _check(lib().atlg_dpotrs_batched(len(kn), pn, pa, pl, pb, plb, nrhs, _stream(stream)),
           "atlg_dpotrs_batched")


def dposv_batched(n, A, lda, B, ldb, nrhs, info, stream=None):
    (kn, pn), (ka, pa), (kl, pl), (kb, pb), (klb, plb) = (_ints(n), _ptrs(A), _ints(lda),
                                                          _ptrs(B), _ints(ldb))
    _check(lib().atlg_dposv_batched(len(kn), pn, pa, pl, pb, plb, nrhs, _ptr(info),
                                    _stream(stream)), "atlg_dposv_batched")


def spotrf_batched(n, A, lda, info, stream=None):
    (kn, pn), (ka, pa), (kl, pl) = _ints(n), _ptrs(A), _ints(lda)
    _check(lib().atlg_spotrf_batched(len(kn), pn, pa, pl, _ptr(info), _stream(stream)),
           "atlg_spotrf_batched")


def sposv_batched(n, A, lda, B, ldb, nrhs, info, stream=None):
    (kn, pn), (ka, pa), (kl, pl), (kb, pb), (klb, plb) = (_ints(n), _ptrs(A), _ints(lda),
                                                          _ptrs(B), _ints(ldb))
    _check(lib().atlg_sposv_batched(len(kn), pn, pa, pl, pb, plb, nrhs, _ptr(info),
                                    _stream(stream)), "atlg_sposv_batched")


def dsposv_batched(n, A, lda, B, ldb, X, ldx, nrhs, info, iter, itermax=30, stream=None):
    (kn, pn), (ka, pa), (kl, pl), (kb, pb), (klb, plb), (kx, px), (klx, plx) = (
        _ints(n), _ptrs(A), _ints(lda), _ptrs(B), _ints(ldb), _ptrs(X), _ints(ldx))
    _check(lib().atlg_dsposv_batched(len(kn), pn, pa, pl, pb, plb, px, plx, nrhs, itermax,
                                     _ptr(info), _ptr(iter), _stream(stream)),
           "atlg_dsposv_batched")


def batched_tune(prec):
    t, w = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().atlg_batched_tune_get(prec, ctypes.byref(t), ctypes.byref(w)),
           "atlg_batched_tune_get")
    return t.value, w.value


def set_batched_tune(prec, t16_max, wg_per_cu):
    _check(lib().atlg_batched_tune_set(prec, t16_max, wg_per_cu), "atlg_batched_tune_set")


# ---------------------------------------------------------------- single / mixed precision

def sgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream=None):
    _check(lib().atlg_sgemm(ta, tb, m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c),
                            ldc, _stream(stream)), "atlg_sgemm")


def sgemm_bf16(m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, split=0, stream=None):
    _check(lib().atlg_sgemm_bf16(m, n, k, alpha, _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc,
                                 split, _stream(stream)), "atlg_sgemm_bf16")


def sgetrf(m, n, a, lda, ipiv, info, gemm_prec=0, stream=None, check=False):
    _check(lib().atlg_sgetrf(m, n, _ptr(a), lda, _ptr(ipiv), _ptr(info), gemm_prec,
                             _stream(stream)), "atlg_sgetrf")
    if check:
        check_stream(stream)


def sgetrs(trans, n, nrhs, a, lda, ipiv, b, ldb, stream=None):
    _check(lib().atlg_sgetrs(trans, n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb,
                             _stream(stream)), "atlg_sgetrs")


def dsgesv(n, nrhs, a, lda, ipiv, b, ldb, x, ldx, info, itermax=30, lowprec=0, refine=0,
           stream=None, inner=False):
    """Returns ITER (LAPACK dsgesv meaning); with inner=True, (ITER, GMRES iterations)."""
    it, inn = ctypes.c_int(0), ctypes.c_int(0)
    _check(lib().atlg_dsgesv(n, nrhs, _ptr(a), lda, _ptr(ipiv), _ptr(b), ldb, _ptr(x), ldx,
                             _ptr(info), itermax, lowprec, refine, ctypes.byref(it),
                             ctypes.byref(inn), _stream(stream)), "atlg_dsgesv")
    return (it.value, inn.value) if inner else it.value
