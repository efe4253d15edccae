"""Empirical search for the batched SPD solvers' decision table.

Written in the style of ATLAS's own install-time searches (the LAPACK NB
search of AtlasBase/Clint/atlas-tlp.base:5578-6220, GetMultSampleTimes /
findNB, and the generated-header emitters such as atlas-mmcpy.base:3369):

* every candidate is timed several times and the minimum is kept;
* each timing is appended as one ATLAS-style summary line
  ("ID=7 ROUT=dposv_batched T=16 WPC=5 N=256 NSYS=512 NRHS=4 TIME=... MFLOP=...",
  MFLOP being the rate in MFLOP/s, as in ATLAS's *.sum files) to
  tune/res/<rout>.sum; a later run reads the file and times only what is
  missing, so an interrupted search resumes;
* the search is coarse first, then refined: block-column width per order
  (16 versus 32 on a ladder of orders, then bisection between the last order
  where 16 wins and the first where 32 wins), then workgroups per compute
  unit on a sample of the uneven workload (a coarse ladder, then the
  neighbours of the best);
* the header math-atlas_amd/csrc/atlg_batched_tune.h, which the library
  compiles in as its defaults, is generated from the *.sum files alone: per
  precision the workgroups per compute unit of the fastest uneven-sample
  line, and a decision tree on the system order (ATLAS's generated
  decision code of atlas-tlp.base:5126-5300 decides on problem size the
  same way) whose leaves are the fastest block-column width on each
  interval between measured orders.

Usage (on a machine with the GPU, the library built):
    python math-atlas_amd/tune/batched_search.py [--prec d|s|both]
        [--res DIR] [--out HEADER] [--samples K] [--quick]
    python math-atlas_amd/tune/batched_search.py --emit-only   (no GPU: header from *.sum)
"""
import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
sys.path.insert(0, PKG)

BIG = 1 << 30  # t16_max that forces 16-wide block columns for every order
ROUT = {"d": "dposv_batched", "s": "sposv_batched"}


def flops_posv(n, nrhs):
    """Cholesky plus two triangular solves: n^3/3 + 2 n^2 nrhs."""
    return n ** 3 / 3.0 + 2.0 * n * n * nrhs


class Sum:
    """Append-only summary file, one 'KEY=value ...' line per timing."""

    KEYS = ("T", "WPC", "N", "NSYS", "NRHS")

    def __init__(self, path, rout):
        self.path, self.rout, self.lines = path, rout, []
        if os.path.exists(path):
            for line in open(path):
                d = dict(tok.split("=", 1) for tok in line.split() if "=" in tok)
                if d.get("ROUT") == rout and "TIME" in d:
                    self.lines.append(d)

    @staticmethod
    def key(d):
        return tuple(str(d[k]) for k in Sum.KEYS)

    def find(self, **kw):
        k = self.key({a: str(b) for a, b in kw.items()})
        for d in self.lines:
            if self.key(d) == k:
                return float(d["TIME"])
        return None

    def get(self, fn, mflop, **kw):
        t = self.find(**kw)
        if t is None:
            t = fn()
            d = {k: str(v) for k, v in kw.items()}
            d.update(ID=str(len(self.lines) + 1), ROUT=self.rout, TIME="%.6e" % t,
                     MFLOP="%.1f" % (mflop / t))
            self.lines.append(d)
            with open(self.path, "a") as f:
                f.write(" ".join("%s=%s" % (k, d[k]) for k in
                                 ("ID", "ROUT") + self.KEYS + ("TIME", "MFLOP")) + "\n")
        return t


def _gpu():
    import torch
    import atlas_gpu as G
    return torch, G


class Workload:
    """SPD systems A = B B^T / n + I (B iid normal) packed on the device,
    with nrhs normal right-hand sides; restored before every timing."""

    def __init__(self, sizes, nrhs, dtype, seed):
        torch, _ = _gpu()
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
        self.mflop = sum(flops_posv(n, nrhs) for n in self.n) / 1e6

    def run(self):
        torch, G = _gpu()
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
    _, G = _gpu()
    G.set_batched_tune(prec, t16, wpc)
    work.run()  # warm-up
    return min(work.run() for _ in range(samples))


def search(prec, res, samples, quick):
    """Times the candidates missing from `res` (a Sum); returns nothing:
    the header is emitted from the summary files afterwards."""
    import numpy as np
    torch, G = _gpu()
    dtype = torch.float64 if prec == 0 else torch.float32
    tag = "d" if prec == 0 else "s"
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    _, wpc0 = G.batched_tune(prec)  # the compiled-in default (nothing was set yet)
    nsys = 2 * cus  # one full wave of persistent workgroups, twice
    nrhs = 4

    # -- stage 1: block-column width per order (coarse ladder) -------------
    ladder = [16, 32, 48, 64, 96, 128, 192, 256, 384, 512]
    if quick:
        ladder = [32, 64, 128, 256, 512]

    def t_at(n, width):
        def measure():
            work = Workload([n] * nsys, nrhs, dtype, seed=n)
            return min_time(work, prec, BIG if width == 16 else 0, wpc0, samples)
        return res.get(measure, nsys * flops_posv(n, nrhs) / 1e6, T=width, WPC=wpc0, N=n,
                       NSYS=nsys, NRHS=nrhs)

    wins16 = {n: t_at(n, 16) <= t_at(n, 32) for n in ladder}
    print("[%s] width 16 wins at:" % tag, [n for n in ladder if wins16[n]], flush=True)
    # refine every change of winner between neighbouring ladder orders by
    # bisection (to 8), so the decision tree's boundaries are measured
    if not quick:
        for a, b in zip(ladder, ladder[1:]):
            if wins16[a] == wins16[b]:
                continue
            lo, hi = a, b
            while hi - lo > 8:
                mid = (lo + hi) // 2
                if (t_at(mid, 16) <= t_at(mid, 32)) == wins16[a]:
                    lo = mid
                else:
                    hi = mid

    # -- stage 2: workgroups per CU on the uneven workload sample ----------
    rng = np.random.default_rng(2024)
    sample = rng.integers(16, 513, size=(600 if quick else 2000)).tolist()
    work = Workload(sample, nrhs, dtype, seed=7)

    def t_wpc(w):
        return res.get(lambda: min_time(work, prec, -1, w, samples), work.mflop, T="tree", WPC=w,
                       N="u%d" % len(sample), NSYS=len(sample), NRHS=nrhs)

    coarse = [1, 2, 4, 8]
    best = min(coarse, key=t_wpc)
    for w in (best - 1, best + 1):
        if 1 <= w <= 16:
            t_wpc(w)
    G.set_batched_tune(prec, -1, wpc0)


def decide(lines):
    """From one precision's summary lines: (workgroups per CU, intervals),
    intervals a list of (last order, width) covering every order."""
    occ = [(float(d["TIME"]), int(d["WPC"])) for d in lines if d["N"].startswith("u")]
    wpc = min(occ)[1] if occ else 2
    by_n = {}
    for d in lines:
        if d["N"].startswith("u"):
            continue
        by_n.setdefault(int(d["N"]), {})[int(d["T"])] = float(d["TIME"])
    pts = sorted((n, min(t, key=t.get)) for n, t in by_n.items() if len(t) == 2)
    if not pts:
        return wpc, [(1 << 30, 16)]
    iv = []
    for (n0, w0), (n1, w1) in zip(pts, pts[1:]):
        if w0 != w1:
            iv.append(((n0 + n1) // 2, w0))  # boundary halfway between measured orders
    iv.append((1 << 30, pts[-1][1]))
    return wpc, iv


def tree(iv, lo, hi, ind):
    """Balanced if/else over intervals iv[lo:hi] (C source lines)."""
    if hi - lo == 1:
        return ["%sreturn %d;" % (ind, iv[lo][1])]
    mid = (lo + hi) // 2
    return (["%sif (n <= %d) {" % (ind, iv[mid - 1][0])] + tree(iv, lo, mid, ind + "  ") +
            ["%s} else {" % ind] + tree(iv, mid, hi, ind + "  ") + ["%s}" % ind])


def emit(path, res_dir, device=""):
    """Writes the header from the summary files in res_dir."""
    lines = [
        "// Defaults of the batched SPD solvers, generated by",
        "// math-atlas_amd/tune/batched_search.py from its summary files",
        "// (math-atlas_amd/tune/res/*.sum)%s." % ((" measured on " + device) if device else ""),
        "// Per precision: persistent workgroups per compute unit, and the",
        "// block-column width (16 or 32) as a decision tree on the system order.",
        "#pragma once",
    ]
    for tag in ("d", "s"):
        sm = Sum(os.path.join(res_dir, ROUT[tag] + ".sum"), ROUT[tag])
        wpc, iv = decide(sm.lines)
        lines.append("//")
        lines.append("// %s (%d lines):" % (ROUT[tag], len(sm.lines)))
        for d in sm.lines:
            lines.append("//   " + " ".join("%s=%s" % (k, d[k]) for k in
                                            ("T", "WPC", "N", "NSYS", "TIME", "MFLOP")))
        lines.append("#define ATLG_BATCH_WPC_%s %d" % (tag.upper(), wpc))
        lines.append("static inline int atlg_batched_width_%s(int n) {" % tag)
        lines += tree(iv, 0, len(iv), "  ")
        lines.append("}")
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
    ap.add_argument("--emit-only", action="store_true")
    args = ap.parse_args()
    os.makedirs(args.res, exist_ok=True)
    device = ""
    if not args.emit_only:
        torch, G = _gpu()
        assert torch.cuda.is_available()
        assert G.lib().atlg_init() == 0
        device = torch.cuda.get_device_name(0)
        for tag, prec in (("d", 0), ("s", 1)):
            if args.prec in (tag, "both"):
                search(prec, Sum(os.path.join(args.res, ROUT[tag] + ".sum"), ROUT[tag]),
                       args.samples, args.quick)
    emit(args.out, args.res, device)


if __name__ == "__main__":
    main()
