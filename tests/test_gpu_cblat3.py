"""The netlib CBLAS level-3 test (AtlasBase/Clint/misc/CBLATST/c_dblat3.f,
the driver ATLAS ships for its cblas_ interface) restated in Python over the
oracle's C restatement of its machinery (DBEG, DMAKE, DMMCH).

For each of cblas_dgemm, cblas_dsymm, cblas_dtrmm, cblas_dtrsm, cblas_dsyrk
and cblas_dsyr2k, every combination of c_dblat3.dat's parameters (N in
{0,1,2,3,5,9}, ALPHA in {0,1,0.7}, BETA in {0,1,1.3}, both storage orders,
all side / uplo / trans / diag options) goes through the host C ABI on
DMAKE's inputs, whose unreferenced parts (rows past M, the other triangle, a
unit diagonal) hold the rogue value -1e10.  Each result is checked as the
Fortran checks it (DCHK1..DCHK5): DMMCH's ratio |expected - computed| /
(eps * gauge) must stay below c_dblat3.dat's threshold 16 and never reach
the half-accurate fatal level, and every element the routine must not write
(padding rows or columns, the other triangle of C) must still be rogue.
"""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
P = json.load(open(os.path.join(HERE, "golden", "atlas_test_params.json")))["cblat3"]
THRESH = P["threshold"]
NS, ALF, BET = P["N"], P["alpha"], P["beta"]
NMAX = max(NS) + 1
ROGUE = -1.0e10
TRANS = (O.NoTrans, O.Trans, O.ConjTrans)


@pytest.fixture(scope="module")
def G():
    import torch
    assert torch.cuda.is_available()
    import atlas_gpu
    assert atlas_gpu.lib().atlg_init() == 0
    return atlas_gpu


def _ld(x):
    return max(x, 1) + 1  # c_dblat3: LDA = MA + 1 (MA < NMAX)


def _make(g, typ, rows, cols, uplo=0, diag=0):
    """DMAKE: (A with leading dimension NMAX, AA with rogue padding)."""
    return O.blat_dmake(g, typ, uplo, diag, rows, cols, NMAX, _ld(rows))


def _to_rm(aa, rows, cols):
    """c_dblat3's C wrappers pass row-major operands: AA's rows x cols part
    (rogue entries included) stored row by row with LD = cols + 1."""
    r = np.full((max(rows, 1), _ld(cols)), ROGUE)
    r[:rows, :cols] = aa[:rows, :cols]
    return np.asfortranarray(r.T), _ld(cols)  # memory: row-major r


def _from_rm(rt, rows, cols):
    return np.asfortranarray(rt.T[:rows, :cols])


def _untouched(arr, before, rows, cols, order):
    """Everything outside the rows x cols operand kept its input value."""
    keep = np.ones(arr.shape, bool)
    if order == "col":
        keep[:rows, :cols] = False
    else:
        keep.T[:rows, :cols] = False  # arr.T is the row-major matrix
    return np.array_equal(arr[keep], before[keep])


class Check:
    def __init__(self):
        self.errmax = 0.0
        self.n = 0

    def mmch(self, ta, tb, m, n, k, alpha, a, b, beta, c, cc):
        f = np.asfortranarray
        fatal, _, emax = O.blat_dmmch(ta, tb, m, n, k, alpha, f(a), f(b), beta, f(c), f(cc))
        assert not fatal, (m, n, k, alpha, beta, emax)
        self.errmax = max(self.errmax, emax)
        self.n += 1


def _call(G, order, fn, dims, ops, tail):
    """Run fn on the column-major operands or on their row-major copies.
    ops: [(aa, rows, cols, written)], returns the written operand's result
    as a rows x cols Fortran array and whether its padding is unchanged."""
    args, out = [], None
    for aa, rows, cols, written in ops:
        if order == "col":
            arr, ld = np.asfortranarray(aa.copy()), aa.shape[0]
        else:
            arr, ld = _to_rm(aa, rows, cols)
        args.append((arr, ld))
        if written:
            out = (arr, arr.copy(), rows, cols)
    fn(G.ColMajor if order == "col" else G.RowMajor, *dims, *tail(args))
    arr, before, rows, cols = out
    res = arr[:rows, :cols] if order == "col" else _from_rm(arr, rows, cols)
    return res, _untouched(arr, before, rows, cols, order)


def _other_triangle(n, uplo):
    return (np.triu(np.ones((n, n), bool), 1) if uplo == O.Lower
            else np.tril(np.ones((n, n), bool), -1))


@pytest.mark.parametrize("order", ["col", "row"])
def test_cblat3_dgemm(G, order):
    # DCHK1 (c_dblat3.f:418-740)
    g, chk = O.BlatGen(), Check()
    for m in NS:
        for n in NS:
            for k in NS:
                for ta in TRANS:
                    for tb in TRANS:
                        ma, na = (m, k) if ta == O.NoTrans else (k, m)
                        mb, nb = (k, n) if tb == O.NoTrans else (n, k)
                        for alpha in ALF:
                            for beta in BET:
                                a, aa = _make(g, O.BLAT_GE, ma, na)
                                b, bb = _make(g, O.BLAT_GE, mb, nb)
                                c, cc = _make(g, O.BLAT_GE, m, n)
                                res, pad = _call(
                                    G, order, G.cblas_dgemm, (ta, tb, m, n, k, alpha),
                                    [(aa, ma, na, False), (bb, mb, nb, False), (cc, m, n, True)],
                                    lambda x: (x[0][0], x[0][1], x[1][0], x[1][1], beta,
                                               x[2][0], x[2][1]))
                                assert pad, ("padding of C written", m, n, k)
                                if m and n:
                                    chk.mmch(ta, tb, m, n, k, alpha, a[:max(ma, 1)],
                                             b[:max(mb, 1)], beta, c[:m], res)
    print("cblas_dgemm %s: %d results, largest ratio %.3g" % (order, chk.n, chk.errmax))
    assert chk.errmax < THRESH


@pytest.mark.parametrize("order", ["col", "row"])
def test_cblat3_dsymm(G, order):
    # DCHK2 (c_dblat3.f:742-1050)
    g, chk = O.BlatGen(), Check()
    for m in NS:
        for n in NS:
            for side in (O.Left, O.Right):
                na = m if side == O.Left else n
                for uplo in (O.Upper, O.Lower):
                    for alpha in ALF:
                        for beta in BET:
                            a, aa = _make(g, O.BLAT_SY, na, na, uplo)
                            b, bb = _make(g, O.BLAT_GE, m, n)
                            c, cc = _make(g, O.BLAT_GE, m, n)
                            res, pad = _call(
                                G, order, G.cblas_dsymm, (side, uplo, m, n, alpha),
                                [(aa, na, na, False), (bb, m, n, False), (cc, m, n, True)],
                                lambda x: (x[0][0], x[0][1], x[1][0], x[1][1], beta,
                                           x[2][0], x[2][1]))
                            assert pad, (order, m, n, side, uplo)
                            if not (m and n):
                                continue
                            if side == O.Left:
                                chk.mmch(O.NoTrans, O.NoTrans, m, n, m, alpha, a[:m, :m],
                                         b[:m, :n], beta, c[:m, :n], res)
                            else:
                                chk.mmch(O.NoTrans, O.NoTrans, m, n, n, alpha, b[:m, :n],
                                         a[:n, :n], beta, c[:m, :n], res)
    print("cblas_dsymm %s: %d results, largest ratio %.3g" % (order, chk.n, chk.errmax))
    assert chk.errmax < THRESH


@pytest.mark.parametrize("rout", ["trmm", "trsm"])
@pytest.mark.parametrize("order", ["col", "row"])
def test_cblat3_dtrmm_dtrsm(G, rout, order):
    # DCHK3 (c_dblat3.f:1052-1408): TRMM checked as alpha op(A) B, TRSM by
    # multiplying the computed X back: op(A) X against alpha B
    g, chk = O.BlatGen(), Check()
    fn = G.cblas_dtrmm if rout == "trmm" else G.cblas_dtrsm
    for m in NS:
        for n in NS:
            for side in (O.Left, O.Right):
                na = m if side == O.Left else n
                for uplo in (O.Upper, O.Lower):
                    for ta in TRANS:
                        for diag in (O.NonUnit, O.Unit):
                            for alpha in ALF:
                                a, aa = _make(g, O.BLAT_TR, na, na, uplo, diag)
                                b, bb = _make(g, O.BLAT_GE, m, n)
                                res, pad = _call(
                                    G, order, fn, (side, uplo, ta, diag, m, n, alpha),
                                    [(aa, na, na, False), (bb, m, n, True)],
                                    lambda x: (x[0][0], x[0][1], x[1][0], x[1][1]))
                                assert pad
                                if not (m and n):
                                    continue
                                zero = np.zeros((m, n), order="F")
                                if rout == "trmm":
                                    if side == O.Left:
                                        chk.mmch(ta, O.NoTrans, m, n, m, alpha, a[:m, :m],
                                                 b[:m, :n], 0.0, zero, res)
                                    else:
                                        chk.mmch(O.NoTrans, ta, m, n, n, alpha, b[:m, :n],
                                                 a[:n, :n], 0.0, zero, res)
                                else:
                                    ab = alpha * b[:m, :n]
                                    if side == O.Left:
                                        chk.mmch(ta, O.NoTrans, m, n, m, 1.0, a[:m, :m], res,
                                                 0.0, zero, ab)
                                    else:
                                        chk.mmch(O.NoTrans, ta, m, n, n, 1.0, res, a[:n, :n],
                                                 0.0, zero, ab)
    print("cblas_d%s %s: %d results, largest ratio %.3g" % (rout, order, chk.n, chk.errmax))
    assert chk.errmax < THRESH


@pytest.mark.parametrize("rout", ["syrk", "syr2k"])
@pytest.mark.parametrize("order", ["col", "row"])
def test_cblat3_dsyrk_dsyr2k(G, rout, order):
    # DCHK4 / DCHK5 (c_dblat3.f:1410-1770, 1772-2190): column j's part of
    # the uplo triangle is checked by DMMCH; for SYR2K the stacked operand
    # [A B] times [B(j,:) A(j,:)] gives (A B^T + B A^T)(:, j) with inner
    # dimension 2K
    g, chk = O.BlatGen(), Check()
    for n in NS:
        for k in NS:
            for uplo in (O.Upper, O.Lower):
                for trans in TRANS:
                    ma, na = (n, k) if trans == O.NoTrans else (k, n)
                    for alpha in ALF:
                        for beta in BET:
                            a, aa = _make(g, O.BLAT_GE, ma, na)
                            b, bb = _make(g, O.BLAT_GE, ma, na)
                            c, cc = _make(g, O.BLAT_SY, n, n, uplo)
                            if rout == "syrk":
                                res, pad = _call(
                                    G, order, G.cblas_dsyrk, (uplo, trans, n, k, alpha),
                                    [(aa, ma, na, False), (cc, n, n, True)],
                                    lambda x: (x[0][0], x[0][1], beta, x[1][0], x[1][1]))
                            else:
                                res, pad = _call(
                                    G, order, G.cblas_dsyr2k, (uplo, trans, n, k, alpha),
                                    [(aa, ma, na, False), (bb, ma, na, False), (cc, n, n, True)],
                                    lambda x: (x[0][0], x[0][1], x[1][0], x[1][1], beta,
                                               x[2][0], x[2][1]))
                            assert pad, (order, m, n, side, uplo)
                            if not n:
                                continue
                            assert np.all(res[_other_triangle(n, uplo)] == ROGUE)
                            nt = trans == O.NoTrans
                            for j in range(n):
                                jj, lj = (0, j + 1) if uplo == O.Upper else (j, n - j)
                                cj, rj = c[jj:jj + lj, j:j + 1], res[jj:jj + lj, j:j + 1]
                                if k == 0:  # op(A) has no columns: C = beta C
                                    chk.mmch(O.NoTrans, O.NoTrans, lj, 1, 0, alpha,
                                             np.zeros((lj, 1)), np.zeros((1, 1)), beta, cj, rj)
                                    continue
                                if rout == "syrk":
                                    x = a[jj:jj + lj, :k] if nt else a[:k, jj:jj + lj]
                                    w = a[j:j + 1, :k].T if nt else a[:k, j:j + 1]
                                    kk = k
                                else:
                                    if nt:
                                        x = np.hstack([a[jj:jj + lj, :k], b[jj:jj + lj, :k]])
                                        w = np.concatenate([b[j, :k], a[j, :k]])[:, None]
                                    else:
                                        x = np.vstack([a[:k, jj:jj + lj], b[:k, jj:jj + lj]])
                                        w = np.concatenate([b[:k, j], a[:k, j]])[:, None]
                                    kk = 2 * k
                                chk.mmch(O.NoTrans if nt else O.Trans, O.NoTrans, lj, 1, kk,
                                         alpha, x, w, beta, cj, rj)
    print("cblas_d%s %s: %d results, largest ratio %.3g" % (rout, order, chk.n, chk.errmax))
    assert chk.errmax < THRESH
