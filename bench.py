"""Benchmark of the FP64 level-3 hot path on one MI355X.

Headline (one JSON line): DGEMM C = A*B + C, column-major FP64,
M = N = K = 16384, alpha = beta = 1, A/B/C iid U[-1,1] (seeded, generated on
the device) -- the N=16384 DGEMM of the FP64 level-3 sweep.  One "step" is
one DGEMM through the C ABI (atlg_dgemm) on torch's current stream; inputs
are resident in HBM before timing starts.

Also reported in the same line (`sweep`): DGEMM N=4096/8192, tall-skinny
M=65536 N=K=256, DTRSM L/L/N/U N=16384 with 16384 RHS, DGETRF and DPOTRF
N=16384 (TFLOP/s with ATLAS's flop formulas); the batched uneven SPD solves
(10000 systems, n ~ U{16..512}, 4 RHS, FP64 and FP32 + FP64 refinement); and
the differentiable Cholesky log-likelihood at N=8192 (forward, and forward +
backward with the factor stored or recomputed).

`roofline` is computed from HIP events recorded on the stream the DGEMM
kernels run on; `cpu_baseline` times the CPU oracle's textbook DGEMM on a
bounded column slab of the same N=16384 problem.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "math-atlas_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

METRIC = "DGEMM TFLOP/s %FP64-peak; FP64-accurate solve time, IR-mixed vs plain dgesv"
PEAK_FP64_MFMA = 78.6      # TFLOP/s, AMD published dense FP64 matrix peak
PEAK_HBM = 8000.0          # GB/s


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--no-sweep", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-cols", type=int, default=160, help="CPU baseline columns per thread")
    ap.add_argument("--only", default="", help="profile helper: run only this sweep item")
    ap.add_argument("--ir-n", type=int, default=32768, help="order of the mixed-precision solve")
    ap.add_argument("--copy-calib", action="store_true",
                    help="profile helper: three 1 GiB device-to-device copies (HBM counter scale)")
    return ap.parse_args()


def uniform(torch, shape, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    t = torch.empty(shape, dtype=torch.float64, device="cuda")
    t.uniform_(-1.0, 1.0, generator=g)
    return t


def timed(torch, fn, reps, warm=1):
    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    return best


def sweep(torch, G, only="", ir_n=32768):
    out = {}

    def want(k):
        return not only or only == k

    for n in (4096, 8192):
        if not want("dgemm_%d" % n):
            continue
        a, b, c = (uniform(torch, (n, n), 10 + i) for i in range(3))
        ms = timed(torch, lambda: G.dgemm(G.NoTrans, G.NoTrans, n, n, n, 1.0, a, n, b, n, 1.0, c, n), 3)
        out["dgemm_%d_tflops" % n] = 2.0 * n ** 3 / ms / 1e9
        del a, b, c
    if want("dgemm_ts"):
        m, n, k = 65536, 256, 256
        a, b, c = uniform(torch, (k, m), 20), uniform(torch, (n, k), 21), uniform(torch, (n, m), 22)
        ms = timed(torch, lambda: G.dgemm(G.NoTrans, G.NoTrans, m, n, k, 1.0, a, m, b, k, 1.0, c, m), 5)
        out["dgemm_ts_65536x256x256_tflops"] = 2.0 * m * n * k / ms / 1e9
        out["dgemm_ts_65536x256x256_GBps"] = 8.0 * (m * k + k * n + 2 * m * n) / ms / 1e6
        del a, b, c
    N = 16384
    if want("dtrsm"):
        a = uniform(torch, (N, N), 30)
        a.mul_(1.0 / N)   # trddom: strict lower part scaled by 1/N (unit diagonal not read)
        b = uniform(torch, (N, N), 31)
        b0 = b.clone()

        def run_trsm():
            b.copy_(b0)
            G.dtrsm(G.Left, G.Lower, G.NoTrans, G.Unit, N, N, 1.0, a, N, b, N)
        ms_copy = timed(torch, lambda: b.copy_(b0), 2)
        ms = timed(torch, run_trsm, 2) - ms_copy
        out["dtrsm_LLNU_16384x16384_tflops"] = float(N) * N * N / ms / 1e9
        del a, b, b0
    if want("laswp"):
        # ATL_dlaswp over all N columns with the interchanges of a partial-
        # pivoting LU of a U[-1,1] matrix (taken from the DGETRF below on a
        # smaller order would change the pattern; a seeded random pivot row
        # per step in [i, N) has the same "row i receives row p" shape).
        import numpy as np
        rng = np.random.default_rng(60)
        piv = np.array([rng.integers(i, N) for i in range(N)], dtype=np.int32)
        ip = torch.from_numpy(piv).to("cuda")
        a = uniform(torch, (N, N), 61)
        ms = timed(torch, lambda: G.dlaswp(N, a, N, 0, N, ip, 1), 3)
        # rows whose content changes, each read and written once per column
        perm = np.arange(N)
        for i in range(N):
            p = piv[i]
            perm[i], perm[p] = perm[p], perm[i]
        moved = int((perm != np.arange(N)).sum())
        out["laswp_16384x16384_ms"] = ms
        out["laswp_16384x16384_rows_moved"] = moved
        out["laswp_16384x16384_GBps"] = 2.0 * 8 * moved * N / ms / 1e6
        del a
    if want("dgetrf"):
        a0 = uniform(torch, (N, N), 40)
        a = a0.clone()
        ipiv = torch.zeros(N, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")

        def run_lu():
            a.copy_(a0)
            G.dgetrf(N, N, a, N, ipiv, info)
        ms_copy = timed(torch, lambda: a.copy_(a0), 2)
        ms = timed(torch, run_lu, 1) - ms_copy
        flops = 2.0 * N ** 3 / 3 - N ** 2 / 2.0 - N / 6.0
        out["dgetrf_16384_tflops"] = flops / ms / 1e9
        out["dgetrf_16384_ms"] = ms
        out["dgetrf_16384_info"] = int(info.item())
        del a, a0
    if want("dpotrf"):
        bm = torch.empty((N, N), dtype=torch.float64, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(50)
        bm.normal_(generator=g)
        s0 = torch.mm(bm, bm.t()).div_(N)   # B B^T / N + N I (symmetric, SPD)
        s0.diagonal().add_(float(N))
        del bm
        s = s0.clone()
        info = torch.zeros(1, dtype=torch.int32, device="cuda")

        def run_chol():
            s.copy_(s0)
            G.dpotrf(G.Lower, N, s, N, info)
        ms_copy = timed(torch, lambda: s.copy_(s0), 2)
        ms = timed(torch, run_chol, 2) - ms_copy
        flops = N ** 3 / 3.0 + N ** 2 / 2.0 + N / 6.0
        out["dpotrf_16384_tflops"] = flops / ms / 1e9
        out["dpotrf_16384_ms"] = ms
        out["dpotrf_16384_info"] = int(info.item())
        del s, s0
    if want("inverses"):
        # trtri / lauum / potri at 8192 on the lower Cholesky factor of
        # B B^T / n + I (B iid normal), each from a fresh copy of the factor
        n = 8192
        bm = torch.empty((n, n), dtype=torch.float64, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(51)
        bm.normal_(generator=g)
        f0 = torch.mm(bm, bm.t()).div_(n)
        f0.diagonal().add_(1.0)
        del bm
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        G.dpotrf(G.Lower, n, f0, n, info)
        w = f0.clone()
        ms_copy = timed(torch, lambda: w.copy_(f0), 2)
        tri = n ** 3 / 3.0
        for name, fn, fl in (
                ("dtrtri", lambda: G.dtrtri(G.Lower, G.NonUnit, n, w, n, info), tri),
                ("dlauum", lambda: G.dlauum(G.Lower, n, w, n), tri),
                ("dpotri", lambda: G.dpotri(G.Lower, n, w, n, info), 2 * tri)):
            def run():
                w.copy_(f0)
                fn()
            ms = timed(torch, run, 2) - ms_copy
            out["%s_8192_ms" % name] = ms
            out["%s_8192_tflops" % name] = fl / ms / 1e9
        del w, f0
    # the extensions report an error string instead of ending the run
    for key, fn in (("config1", lambda: bench_config1(G)),
                    ("ir", lambda: bench_ir(torch, G, ir_n)),
                    ("batched", lambda: bench_batched(torch, G)),
                    ("cholgrad", lambda: bench_cholgrad(torch))):
        if not want(key):
            continue
        try:
            out.update(fn())
        except Exception as e:  # noqa: BLE001
            out["%s_error" % key] = "%s: %s" % (type(e).__name__, e)
        torch.cuda.empty_cache()
    torch.cuda.empty_cache()
    return out


def haar(torch, n, g):
    """Haar-random orthogonal n x n (QR of a Gaussian, R's diagonal signs
    folded into Q)."""
    q, r = torch.linalg.qr(torch.randn((n, n), generator=g, dtype=torch.float64, device="cuda"))
    q.mul_(torch.sign(torch.diagonal(r)).unsqueeze(0))
    del r
    return q


def bench_ir(torch, G, n):
    """Mixed-precision dgesv at order n, one RHS b iid normal(0,1),
    A = U diag(s) V^T with U, V Haar orthogonal and s_i = kappa^(-i/(n-1)),
    kappa in {1e2, 1e6, 1e10}: plain FP64 dgesv (getrf + getrs) against
    atlg_dsgesv with FP32, BF16 and split-BF16 factors, each with classical
    refinement and with GMRES-IR, and each solution's normwise backward
    error ||b - A x||_inf / (||A||_inf ||x||_inf + ||b||_inf)."""
    import time as _t
    g = torch.Generator(device="cuda")
    g.manual_seed(32768)
    t0 = _t.time()
    U = haar(torch, n, g)
    V = haar(torch, n, g)
    out = {"ir_n": n, "ir_setup_s": 0.0}
    b = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
    info = torch.zeros(1, dtype=torch.int32, device="cuda")
    x = torch.zeros(n, dtype=torch.float64, device="cuda")
    ar = torch.arange(n, dtype=torch.float64, device="cuda")
    for kappa in (1e2, 1e6, 1e10):
        sv = torch.pow(torch.tensor(kappa, dtype=torch.float64, device="cuda"), -ar / (n - 1))
        # column-major A = row-major A^T = V diag(s) U^T
        Acm = torch.mm(V * sv.unsqueeze(0), U.t().contiguous())
        A = Acm.clone()
        anrm = float(Acm.abs().sum(dim=0).max())  # ||A||_inf: column sums of A^T
        tag = "%.0e" % kappa

        ax = torch.empty_like(x)

        def berr():
            # A x through this library's DGEMM on the column-major buffer of A
            G.dgemm(G.NoTrans, G.NoTrans, n, 1, n, 1.0, Acm, n, x, n, 0.0, ax, n)
            r = b - ax
            rn, xn, bn = float(r.abs().max()), float(x.abs().max()), float(b.abs().max())
            return rn / (anrm * xn + bn), rn, xn

        def run64():
            A.copy_(Acm)
            x.copy_(b)
            G.dgetrf(n, n, A, n, ipiv, info)
            G.dgetrs(G.NoTrans, n, 1, A, n, ipiv, x, n)

        its = {}

        def runmix(lp, rf):
            A.copy_(Acm)
            its[lp] = G.dsgesv(n, 1, A, n, ipiv, b, n, x, n, info, 30, lp, refine=rf, inner=True)

        ms_copy = timed(torch, lambda: A.copy_(Acm), 2)
        if kappa == 1e2:
            out["ir_setup_s"] = round(_t.time() - t0, 1)
            ms64 = timed(torch, run64, 1) - ms_copy
            out["ir_dgesv_fp64_ms"] = ms64
            out["ir_dgesv_fp64_berr"], out["ir_dgesv_fp64_rnorm"], out["ir_dgesv_fp64_xnorm"] = berr()
            out["ir_anorm"] = anrm
            out["ir_dgetrf_fp64_tflops"] = (2.0 * n ** 3 / 3) / ms64 / 1e9
        # factor precision x refinement; the single-BF16 factor only with
        # GMRES-IR (classical refinement cannot converge from its 2^-8
        # roundoff at this order)
        best = None
        for lp, name in ((0, "fp32"), (2, "bf16x6"), (3, "bf16x3"), (1, "bf16")):
            for rf, rname in ((0, "ir"), (1, "gmres")):
                if lp == 1 and rf == 0:
                    continue
                ms = timed(torch, lambda: runmix(lp, rf), 1) - ms_copy
                key = "ir_k%s_%s_%s" % (tag, name, rname)
                out[key + "_ms"] = ms
                out[key + "_iter"], out[key + "_inner"] = its[lp]
                out[key + "_berr"] = berr()[0]
                out[key + "_speedup"] = out["ir_dgesv_fp64_ms"] / ms
                if its[lp][0] >= 0 and (best is None or ms < best[1]):
                    best = (key, ms)
        if best is not None:
            out["ir_k%s_best" % tag] = best[0]
            out["ir_k%s_best_speedup" % tag] = out["ir_dgesv_fp64_ms"] / best[1]
        del Acm, A
    del U, V
    return out


def batched_workload(torch, nsys=10000, lo=16, hi=512, nrhs=4, seed=4):
    """nsys SPD systems, n ~ U{lo..hi} (seeded), A = B B^T / n + I with B iid
    normal(0,1), packed column-major (lda = n) by order; nrhs normal
    right-hand sides per system."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = rng.integers(lo, hi + 1, size=nsys)
    order = np.argsort(ns, kind="stable")
    offs = np.zeros(nsys, dtype=np.int64)
    tot = 0
    for k in order:
        offs[k] = tot
        tot += int(ns[k]) ** 2
    A0 = torch.empty(tot, dtype=torch.float64, device="cuda")
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    for n in np.unique(ns):
        idx = np.nonzero(ns == n)[0]
        n = int(n)
        o = int(offs[idx].min())
        b = torch.randn((len(idx), n, n), generator=g, dtype=torch.float64, device="cuda")
        a = torch.bmm(b, b.transpose(1, 2).contiguous()).div_(n)
        a.diagonal(dim1=1, dim2=2).add_(1.0)
        A0[o:o + len(idx) * n * n] = a.reshape(-1)
    boffs = np.cumsum(np.concatenate([[0], ns[:-1] * nrhs])).astype(np.int64)
    B0 = torch.randn(int(ns.sum()) * nrhs, generator=g, dtype=torch.float64, device="cuda")
    return [int(x) for x in ns], offs, A0, boffs, B0


def bench_batched(torch, G):
    """Batched uneven SPD solves: 10000 systems, n ~ U{16..512}, 4 RHS,
    potrf + potrs in one persistent launch (FP64), and FP32 factor + FP64
    refinement (atlg_dsposv_batched)."""
    import numpy as np
    nrhs = 4
    ns, offs, A0, boffs, B0 = batched_workload(torch, nrhs=nrhs)
    A, B = A0.clone(), B0.clone()
    X = torch.zeros_like(B0)
    pa = [A.data_ptr() + 8 * int(o) for o in offs]
    pb = [B.data_ptr() + 8 * int(o) for o in boffs]
    px = [X.data_ptr() + 8 * int(o) for o in boffs]
    nsys = len(ns)
    info = torch.zeros(nsys, dtype=torch.int32, device="cuda")
    it = torch.zeros(nsys, dtype=torch.int32, device="cuda")

    def restore():
        A.copy_(A0)
        B.copy_(B0)

    def run64():
        restore()
        G.dposv_batched(ns, pa, ns, pb, ns, nrhs, info)

    def runmix():
        restore()
        G.dsposv_batched(ns, pa, ns, pb, ns, px, ns, nrhs, info, it)

    ms_copy = timed(torch, restore, 3)
    ms64 = timed(torch, run64, 3) - ms_copy
    bad64 = int((info != 0).sum().item())
    msmix = timed(torch, runmix, 3) - ms_copy
    badmix = int((info != 0).sum().item())
    n = np.asarray(ns, dtype=np.float64)
    flops = float((n ** 3 / 3 + n ** 2 / 2 + n / 6).sum() + (2 * n ** 2 * nrhs).sum())
    byts = float((8 * n ** 2 + 4 * n ** 2 + 16 * n * nrhs).sum())
    its = it.cpu().numpy()
    out = {
        "batched_posv_10000_ms": ms64,
        "batched_posv_10000_systems_per_s": nsys / ms64 * 1e3,
        "batched_posv_10000_tflops": flops / ms64 / 1e9,
        "batched_posv_10000_GBps": byts / ms64 / 1e6,
        "batched_posv_10000_nonzero_info": bad64,
        "batched_dsposv_10000_ms": msmix,
        "batched_dsposv_10000_fallbacks": int((its < 0).sum()),
        "batched_dsposv_10000_max_iter": int(its.max()) if len(its) else 0,
        "batched_dsposv_10000_nonzero_info": badmix,
    }
    # backward error of the refined solutions on a few systems (FP64 check)
    worst = 0.0
    for k in range(0, nsys, 997):
        m = ns[k]
        a = A0[int(offs[k]):int(offs[k]) + m * m].view(m, m)
        x = X[int(boffs[k]):int(boffs[k]) + m * nrhs].view(nrhs, m).t()
        b = B0[int(boffs[k]):int(boffs[k]) + m * nrhs].view(nrhs, m).t()
        r = (b - a @ x).abs().amax(dim=0)
        den = a.abs().sum(dim=1).amax() * x.abs().amax(dim=0) + b.abs().amax(dim=0)
        worst = max(worst, float((r / den).max()))
    out["batched_dsposv_10000_sample_backward_error"] = worst
    del A, B, X, A0, B0
    return out


def bench_cholgrad(torch, n=8192):
    """Differentiable Cholesky log-likelihood: RBF kernel (length-scale 0.2)
    over n points iid U[0,1]^3 plus 1e-2 I, y iid normal; forward and
    forward + backward through the hand-written adjoint kernels."""
    from atlas_gpu import autograd as AG
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.rand((n, 3), generator=g, dtype=torch.float64, device="cuda")
    d2 = torch.cdist(x, x).pow_(2)
    K0 = torch.exp(-d2 / (2 * 0.2 ** 2))
    K0.diagonal().add_(1e-2)
    del d2
    y = torch.randn(n, generator=g, dtype=torch.float64, device="cuda")
    out = {}
    ms_f = timed(torch, lambda: AG.gaussian_loglik(K0, y, store_factor=True), 3)
    out["cholgrad_8192_forward_ms"] = ms_f
    out["cholgrad_8192_forward_potrf_tflops"] = (n ** 3 / 3.0) / ms_f / 1e9
    for store in (True, False):
        K = K0.clone().requires_grad_(True)
        yy = y.clone().requires_grad_(True)

        def fb():
            K.grad = None
            yy.grad = None
            AG.gaussian_loglik(K, yy, store_factor=store).backward()
        out["cholgrad_8192_fwd_bwd_ms_%s" % ("stored" if store else "recomputed")] = \
            timed(torch, fb, 3)
        del K, yy
    del K0
    return out


def host_cores():
    """Threads the CPU legs use: the host's cores, capped by OMP_NUM_THREADS
    when the environment sets it (the CPU share a GPU box grants)."""
    n = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return n


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, ncols_per_thread):
    """The oracle's textbook DGEMM (ATL_drefgemm loop order) on a column slab
    of the same N=16384 problem, split over the host's cores."""
    import numpy as np
    import oracle as O
    th = host_cores()
    ncols = ncols_per_thread * th
    rng = np.random.default_rng(7)
    a = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n)))
    b = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, ncols)))
    c = np.asfortranarray(rng.uniform(-1.0, 1.0, (n, ncols)))
    sec = O.time_refgemm_slab_mt(n, ncols, n, a, b, c, th)
    flops = 2.0 * n * ncols * n
    return {"value": flops / sec / 1e12, "unit": "TFLOP/s", "cores": th, "kind": "port",
            "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": "oracle textbook DGEMM (ATL_drefgemmNN loop order), M=K=%d, first %d "
                      "columns of C (%.1f GFLOP in %.1f s), %d threads" % (
                          n, ncols, flops / 1e9, sec, th)}


def bench_config1(G):
    """Config 1: cblas_dgemm M=N=K=1000, alpha=beta=1, U[-1,1], through the
    host-pointer C ABI (staging to and from HBM included), checked with
    l3blastst's residual; beside it the oracle's textbook DGEMM of the same
    problem on the host's cores."""
    import numpy as np
    import oracle as O
    n = 1000
    rng = np.random.default_rng(1)
    a, b, c0 = (np.asfortranarray(rng.uniform(-1.0, 1.0, (n, n))) for _ in range(3))
    c = c0.copy(order="F")
    G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, n, n, n, 1.0, a, n, b, n, 1.0, c, n)
    best = None
    for _ in range(5):
        c = c0.copy(order="F")
        t0 = time.perf_counter()
        G.cblas_dgemm(G.ColMajor, G.NoTrans, G.NoTrans, n, n, n, 1.0, a, n, b, n, 1.0, c, n)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    ref = c0.copy(order="F")
    th = host_cores()
    cpu = min(O.time_refgemm_slab_mt(n, n, n, a, b, ref.copy(order="F"), th) for _ in range(3))
    O.refgemm(O.NoTrans, O.NoTrans, n, n, n, 1.0, a, b, 1.0, ref)
    ratio = O.gemm_resid(O.NoTrans, O.NoTrans, n, n, n, 1.0, a, b, 1.0, c0, c, ref)
    flops = 2.0 * n ** 3
    return {"config1_cblas_dgemm_1000_ms": best * 1e3,
            "config1_cblas_dgemm_1000_gflops": flops / best / 1e9,
            "config1_l3blastst_ratio": ratio,
            "config1_cpu_oracle_ms": cpu * 1e3, "config1_cpu_oracle_threads": th}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the level-3 path on a single "
              "MI355X; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import atlas_gpu as G

    assert torch.cuda.is_available(), "bench.py needs an MI355X"
    torch.cuda.set_device(0)
    G.lib()
    assert G.lib().atlg_init() == 0

    if args.only:
        print(json.dumps(sweep(torch, G, args.only, args.ir_n)), flush=True)
        return
    if args.copy_calib:
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each": 2 * src.numel() * 8, "copies": 3}), flush=True)
        return

    n = args.n
    a = uniform(torch, (n, n), 1)
    b = uniform(torch, (n, n), 2)
    c = uniform(torch, (n, n), 3)
    stream = torch.cuda.current_stream()

    def step():
        G.dgemm(G.NoTrans, G.NoTrans, n, n, n, 1.0, a, n, b, n, 1.0, c, n, stream=stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    # per-launch events on the kernels' stream, plus the whole timed region
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(args.steps)]
    t0 = time.perf_counter()
    for e0, e1 in evs:
        e0.record(stream)
        step()
        e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [e0.elapsed_time(e1) for e0, e1 in evs]
    ms_region = evs[0][0].elapsed_time(evs[-1][1])
    ms_per_step = ms_region / args.steps
    flops = 2.0 * n ** 3
    value = flops / ms_per_step / 1e9
    avg_launch = sum(per) / len(per)
    achieved = flops / avg_launch / 1e9

    # spot check of the result rows against torch's FP64 GEMM (checker only)
    rows = torch.arange(0, n, 4099, device="cuda")
    a2, b2, c2 = uniform(torch, (n, n), 1), uniform(torch, (n, n), 2), uniform(torch, (n, n), 3)
    G.dgemm(G.NoTrans, G.NoTrans, n, n, n, 1.0, a2, n, b2, n, 1.0, c2, n, stream=stream)
    # column-major X is stored as x (cols, ld): X = x^T, so C[rows] = A[rows] B + C0[rows]
    a_rows = a2.t()[rows].contiguous()
    bmat = b2.t().contiguous()
    ref = torch.mm(a_rows, bmat) + uniform(torch, (n, n), 3).t()[rows]
    got = c2.t()[rows]
    rel = float((got - ref).abs().max() / ref.abs().max())
    del a2, b2, c2, ref, got, a_rows, bmat

    traffic = None
    tf = os.path.join(ROOT, "profiles", "dgemm_n16384_traffic.json")
    if os.path.exists(tf):
        with open(tf) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "DGEMM NN column-major FP64 M=N=K=%d alpha=beta=1, U[-1,1]" % n,
                   "M": n, "N": n, "K": n, "alpha": 1.0, "beta": 1.0},
        "pct_fp64_peak": round(100.0 * value / PEAK_FP64_MFMA, 2),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": PEAK_FP64_MFMA,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_MFMA, 4),
                     "traffic": traffic, "kernel": "gemm_nn_f64",
                     "avg_launch_ms": round(avg_launch, 4),
                     # sustained-load drift: per-launch spread and first vs last launches
                     "min_launch_ms": round(min(per), 4), "max_launch_ms": round(max(per), 4),
                     "first5_avg_ms": round(sum(per[:5]) / len(per[:5]), 4),
                     "last5_avg_ms": round(sum(per[-5:]) / len(per[-5:]), 4)},
        "check_max_rel_err": rel,
        "wall_s_timed": round(wall, 3),
    }
    del a, b, c
    torch.cuda.empty_cache()
    if not args.no_sweep:
        # five significant digits: backward errors sit near 1e-15
        out["sweep"] = {k: (float("%.5g" % v) if isinstance(v, float) else v)
                        for k, v in sweep(torch, G, "", args.ir_n).items()}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, args.cpu_cols)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
