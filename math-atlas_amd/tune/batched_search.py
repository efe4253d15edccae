"""Empirical search for the batched SPD solvers' decision table.

Written in the style of ATLAS's own install-time searches (the LAPACK NB
search of AtlasBase/Clint/atlas-tlp.base:5578-6220, GetMultSampleTimes /
findNB, and the generated-header emitters such as atlas-mmcpy.base:3369):

* every candidate is timed several times and the minimum is kept;
* each timing is appended to a results file, and a later run reads the file
  and times only what is missing, so an interrupted search resumes;
* the search is coarse first, then refined: block-column width per order
  (16 versus 32 on a ladder of orders, then bisection between the last order
  where 16 wins and the first where 32 wins), then workgroups per compute
  unit on a sample of the uneven workload (a coarse ladder, then the
  neighbours of the best);
* the winners are written as a C header, math-atlas_amd/csrc/
  atlg_batched_tune.h, that the library compiles in as its defaults.

Usage (on a machine with the GPU, the library built):
    python math-atlas_amd/tune/batched_search.py [--prec d|s|both]
        [--res DIR] [--out HEADER] [--samples K] [--quick]
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, PKG)

import numpy as np  # noqa: E402
import torch  # noqa: E402

import atlas_gpu as G  # noqa: E402

BIG = 1 << 30  # t16_max that forces 16-wide block columns for every order


class Results:
    """Append-only timing cache: one 'key value' line per measurement."""

    def __init__(self, path):
        self.path = path
        self.t = {}
        if os.path.exists(path):
            for line in open(path):
                parts = line.split()
                if len(parts) == 2:
                    self.t[parts[0]] = float(parts[1])

    def get(self, key, fn):
        if key not in self.t:
            v = fn()
            self.t[key] = v
            with open(self.path, "a") as f:
                f.write("%s %.6e\n" % (key, v))
        return self.t[key]


class Workload:
    """SPD systems A = B B^T / n + I (B iid normal) packed on the device,
    with nrhs normal right-hand sides; restored before every timing."""

    def __init__(self, sizes, nrhs, dtype, seed):
        g = torch.Generator(device="cuda").manual_seed(seed)
        self.n = [int(x) for x in sizes]
        self.dtype = dtype
        tot = sum(n * n for n in self.n)
        self.A0 = torch.empty(max(tot, 1), dtype=dtype, device="cuda")
        off, self.aoff = 0, []
        for n in sorted(set(self.n)):
            idx = [k for k, m in enumerate(self.n) if m == n]
            b = torch.randn((len(idx), n, n), generator=g, dtype=torch.float64, device="cuda")
            a = torch.bmm(b, b.transpose(1, 2).contiguous()) / n
            a += torch.eye(n, dtype=torch.float64, device="cuda")
            for j, k in enumerate(idx):
                self.A0[off:off + n * n] = a[j].reshape(-1).to(dtype)
                self.aoff.append((k, off))
                off += n * n
        self.aoff = [o for _, o in sorted(self.aoff)]
        self.A = self.A0.clone()
        self.nrhs = nrhs
        boff, btot = [], 0
        for n in self.n:
            boff.append(btot)
            btot += n * nrhs
        self.B0 = torch.randn(max(btot, 1), generator=g, dtype=torch.float64,
                              device="cuda").to(dtype)
        self.B = self.B0.clone()
        es = self.A.element_size()
        self.pa = [self.A.data_ptr() + o * es for o in self.aoff]
        self.pb = [self.B.data_ptr() + o * es for o in boff]
        self.info = torch.zeros(len(self.n), dtype=torch.int32, device="cuda")

    def run(self):
        self.A.copy_(self.A0)
        self.B.copy_(self.B0)
        fn = G.dposv_batched if self.dtype == torch.float64 else G.sposv_batched
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(s)
        fn(self.n, self.pa, self.n, self.pb, self.n, self.nrhs, self.info)
        e1.record(s)
        e1.synchronize()
        if int(self.info.abs().sum().item()) != 0:
            raise RuntimeError("search workload is not positive definite")
        return e0.elapsed_time(e1) * 1e-3


def min_time(work, prec, t16, wpc, samples):
    G.set_batched_tune(prec, t16, wpc)
    work.run()  # warm-up
    return min(work.run() for _ in range(samples))


def search(prec, res, samples, quick):
    dtype = torch.float64 if prec == 0 else torch.float32
    tag = "d" if prec == 0 else "s"
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    _, wpc0 = G.batched_tune(prec)
    nsys_uniform = 2 * cus  # one full wave of persistent workgroups, twice

    # -- stage 1: block-column width per order (coarse ladder) -------------
    ladder = [16, 32, 48, 64, 96, 128, 192, 256, 384, 512]
    if quick:
        ladder = [32, 64, 128, 256, 512]

    def t_at(n, width):
        key = "%s_width%d_n%d_w%d" % (tag, width, n, wpc0)
        work = None

        def measure():
            nonlocal work
            work = Workload([n] * nsys_uniform, 4, dtype, seed=n)
            return min_time(work, prec, BIG if width == 16 else 0, wpc0, samples)

        return res.get(key, measure)

    wins16 = {n: t_at(n, 16) <= t_at(n, 32) for n in ladder}
    print("[%s] width 16 wins at:" % tag, [n for n in ladder if wins16[n]], flush=True)
    # crossover: last ladder order where 16 wins before 32 wins for good
    last16 = 0
    for n in ladder:
        if wins16[n]:
            last16 = n
    # -- refine: bisection between last16 and the next ladder order -------
    above = [n for n in ladder if n > last16]
    if last16 and above and not quick:
        lo, hi = last16, above[0]
        while hi - lo > 8:
            mid = (lo + hi) // 2
            if t_at(mid, 16) <= t_at(mid, 32):
                lo = mid
            else:
                hi = mid
        last16 = lo
    t16_max = last16
    print("[%s] t16_max = %d" % (tag, t16_max), flush=True)

    # -- stage 2: workgroups per CU on the uneven workload sample ----------
    rng = np.random.default_rng(2024)
    sample = rng.integers(16, 513, size=(600 if quick else 2000)).tolist()
    work = Workload(sample, 4, dtype, seed=7)

    def t_wpc(w):
        key = "%s_uneven%d_t%d_w%d" % (tag, len(sample), t16_max, w)
        return res.get(key, lambda: min_time(work, prec, t16_max, w, samples))

    coarse = [1, 2, 4, 8]
    best = min(coarse, key=t_wpc)
    for w in (best - 1, best + 1):
        if 1 <= w <= 16:
            t_wpc(w)
    cand = [w for w in range(1, 17)
            if ("%s_uneven%d_t%d_w%d" % (tag, len(sample), t16_max, w)) in res.t]
    best = min(cand, key=t_wpc)
    print("[%s] wg_per_cu = %d (%.3f ms on the %d-system sample)"
          % (tag, best, 1e3 * t_wpc(best), len(sample)), flush=True)
    table = [(n, t_at(n, 16), t_at(n, 32)) for n in ladder]
    return t16_max, best, table, [(w, t_wpc(w)) for w in sorted(cand)]


def emit(path, found, device):
    lines = [
        "// Decision table of the batched SPD solvers: {t16_max, wg_per_cu} for FP64",
        "// (D) and FP32 (S).  Generated by math-atlas_amd/tune/batched_search.py on",
        "// %s, %s." % (device, time.strftime("%Y-%m-%d")),
        "//",
    ]
    for tag in ("d", "s"):
        if tag not in found:
            continue
        t16, wpc, table, occ = found[tag]
        lines.append("// %s: order -> seconds with 16-wide / 32-wide block columns" % tag.upper())
        for n, a, b in table:
            lines.append("//   n=%4d  %.4e  %.4e%s" % (n, a, b, "  *16" if a <= b else ""))
        lines.append("// %s: workgroups per CU -> seconds on the uneven sample" % tag.upper())
        for w, t in occ:
            lines.append("//   %2d  %.4e%s" % (w, t, "  *" if w == wpc else ""))
    lines.append("#pragma once")
    d = found.get("d", (96, 2))[:2]
    s = found.get("s", (128, 2))[:2]
    lines.append("#define ATLG_BATCH_TUNE_D {%d, %d}" % d)
    lines.append("#define ATLG_BATCH_TUNE_S {%d, %d}" % s)
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("wrote", path)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--prec", default="both", choices=["d", "s", "both"])
    ap.add_argument("--res", default=os.path.join(HERE, "res"))
    ap.add_argument("--out", default=os.path.join(PKG, "csrc", "atlg_batched_tune.h"))
    ap.add_argument("--samples", type=int, default=3)
    ap.add_argument("--quick", action="store_true")
    args = ap.parse_args()
    os.makedirs(args.res, exist_ok=True)
    assert torch.cuda.is_available()
    assert G.lib().atlg_init() == 0
    found = {}
    for tag, prec in (("d", 0), ("s", 1)):
        if args.prec in (tag, "both"):
            res = Results(os.path.join(args.res, "batched_%s.txt" % tag))
            found[tag] = search(prec, res, args.samples, args.quick)
    emit(args.out, found, torch.cuda.get_device_name(0))


if __name__ == "__main__":
    main()
