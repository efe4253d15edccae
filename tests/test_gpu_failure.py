"""Bounded-wait failures are reported, never returned silently.

The getrf panel (pivot-candidate mailboxes between workgroups) and the
few-RHS triangular solve (per-block ready flags) wait on other workgroups of
their own launch.  Each wait is bounded.  These tests force the bound from
the host (atlg_set_wait_limit(1): give up at the first poll that finds the
data not yet published), so the waits end early and every kernel exits
normally, and check that:

* the host entry points report through the failure hook (the Python
  binding turns it into AtlgWaitError) with the right failure bit;
* getrf leaves pivots in range (identity for the failed panels);
* the device API reports through the stream's failure word and
  atlg_dsgesv returns ATLG_EWAIT;
* with the default bound the same calls succeed and report nothing.
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
    yield atlas_gpu
    atlas_gpu.set_wait_limit(0)


def test_getrf_panel_failure_reported(G):
    n = 2100  # panels of 2100 rows: nine workgroups exchange pivot candidates
    a0 = O.gegen(n, n, 2100 * 2100 + 2100)
    G.set_wait_limit(1)
    try:
        lu = a0.copy(order="F")
        ipiv = np.full(n, -7, dtype=np.int32)
        with pytest.raises(G.AtlgWaitError) as ei:
            G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv)
        assert ei.value.bits & G.FAIL_PANEL
        assert np.all((ipiv >= 0) & (ipiv < n))
    finally:
        G.set_wait_limit(0)
    lu = a0.copy(order="F")
    ipiv = np.zeros(n, dtype=np.int32)
    assert G.clapack_dgetrf(G.ColMajor, n, n, lu, n, ipiv) == 0
    assert O.lu_resid(n, n, a0, lu, ipiv) <= 100


def test_trsv_failure_reported(G):
    n = 8192  # 64 blocks of 128 rows, one right-hand side: the sync-free solve
    rng = np.random.default_rng(1)
    a = np.asfortranarray(np.tril(rng.uniform(-1, 1, (n, n))) / n + 4 * np.eye(n))
    b0 = np.asfortranarray(rng.uniform(-1, 1, (n, 1)))
    G.set_wait_limit(1)
    try:
        b = b0.copy(order="F")
        with pytest.raises(G.AtlgWaitError) as ei:
            G.cblas_dtrsm(G.ColMajor, G.Left, G.Lower, G.NoTrans, G.NonUnit, n, 1, 1.0, a, n, b, n)
        assert ei.value.bits & G.FAIL_TRSV
    finally:
        G.set_wait_limit(0)
    b = b0.copy(order="F")
    G.cblas_dtrsm(G.ColMajor, G.Left, G.Lower, G.NoTrans, G.NonUnit, n, 1, 1.0, a, n, b, n)
    r = np.tril(a) @ b - b0
    assert np.abs(r).max() <= 1e-12 * (np.abs(np.tril(a)).sum(axis=1).max() * np.abs(b).max())


def test_device_api_failure_word(G):
    import torch
    n = 2100
    a0 = O.gegen(n, n, 77)
    G.set_wait_limit(1)
    try:
        da = G.from_fortran(a0)
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        with pytest.raises(G.AtlgWaitError) as ei:
            G.dgetrf(n, n, da, n, ipiv, info, check=True)
        assert ei.value.bits & G.FAIL_PANEL
        assert int(info.item()) == -2 ** 31
        p = ipiv.cpu().numpy()
        assert np.all((p >= 0) & (p < n))
        # dsgesv synchronizes and returns ATLG_EWAIT (its FP32 factor and
        # the FP64 fallback both take the bounded panels)
        m = 2048
        a = torch.randn((m, m), dtype=torch.float64, device="cuda") + m * torch.eye(
            m, dtype=torch.float64, device="cuda")
        b = torch.randn(m, dtype=torch.float64, device="cuda")
        x = torch.zeros(m, dtype=torch.float64, device="cuda")
        piv = torch.zeros(m, dtype=torch.int32, device="cuda")
        with pytest.raises(G.AtlgWaitError):
            G.dsgesv(m, 1, a, m, piv, b, m, x, m, info)
        assert G.stream_status() == 0  # read and cleared by dsgesv
    finally:
        G.set_wait_limit(0)
    da = G.from_fortran(a0)
    G.dgetrf(n, n, da, n, ipiv, info, check=True)
    assert int(info.item()) == 0
    assert O.lu_resid(n, n, a0, G.to_fortran(da), ipiv.cpu().numpy()) <= 100
