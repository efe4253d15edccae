"""The batched solvers' tuning chain without a GPU: summary (*.sum) lines
in, generated header out.  The header's decision tree is compiled with gcc
and evaluated on every order, against the winners the summary lines say."""
import importlib.util
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
TUNE = os.path.join(HERE, "..", "math-atlas_amd", "tune", "batched_search.py")


def _mod():
    spec = importlib.util.spec_from_file_location("batched_search", TUNE)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def _write_sum(path, rout, table, occ):
    with open(path, "w") as f:
        i = 1
        for n, t16, t32 in table:
            for t, v in ((16, t16), (32, t32)):
                f.write("ID=%d ROUT=%s T=%d WPC=4 N=%d NSYS=512 NRHS=4 TIME=%.6e MFLOP=1.0\n"
                        % (i, rout, t, n, v))
                i += 1
        for w, v in occ:
            f.write("ID=%d ROUT=%s T=tree WPC=%d N=u2000 NSYS=2000 NRHS=4 TIME=%.6e MFLOP=1.0\n"
                    % (i, rout, w, v))
            i += 1


def test_sum_to_header(tmp_path):
    m = _mod()
    # 16 wins up to 96 and again at 384 (a non-monotone winner: the tree
    # must keep it), 32 at 128..256 and 512
    d_table = [(16, 1.0, 2.0), (32, 1.0, 2.0), (64, 1.0, 2.0), (96, 1.0, 2.0),
               (128, 2.0, 1.0), (256, 2.0, 1.0), (384, 1.0, 2.0), (512, 2.0, 1.0)]
    s_table = [(16, 1.0, 2.0), (512, 1.0, 2.0)]
    _write_sum(tmp_path / "dposv_batched.sum", "dposv_batched", d_table,
               [(1, 9.0), (2, 5.0), (3, 4.0), (4, 4.5)])
    _write_sum(tmp_path / "sposv_batched.sum", "sposv_batched", s_table, [(2, 3.0), (4, 2.0)])
    hdr = tmp_path / "tune.h"
    m.emit(str(hdr), str(tmp_path))
    text = hdr.read_text()
    assert "#define ATLG_BATCH_WPC_D 3" in text and "#define ATLG_BATCH_WPC_S 4" in text
    prog = tmp_path / "t.c"
    prog.write_text('#include <stdio.h>\n#include "tune.h"\nint main(void) {\n'
                    '  for (int n = 1; n <= 600; n++) printf("%d %d\\n", '
                    'atlg_batched_width_d(n), atlg_batched_width_s(n));\n  return 0;\n}\n')
    exe = tmp_path / "t"
    subprocess.run(["gcc", "-O1", "-o", str(exe), str(prog), "-I", str(tmp_path)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = [tuple(int(x) for x in line.split()) for line in out if line]
    # boundaries halfway between measured orders whose winners differ
    def want_d(n):
        if n <= 112:
            return 16
        if n <= 320:
            return 32
        if n <= 448:
            return 16
        return 32
    for n in range(1, 601):
        assert got[n - 1] == (want_d(n), 16), (n, got[n - 1])
    # every measured order decides as measured
    for n, t16, t32 in d_table:
        assert got[n - 1][0] == (16 if t16 <= t32 else 32)


def test_sum_lines_resume(tmp_path):
    m = _mod()
    path = tmp_path / "dposv_batched.sum"
    sm = m.Sum(str(path), "dposv_batched")
    calls = []
    t = sm.get(lambda: calls.append(1) or 0.5, 100.0, T=16, WPC=2, N=64, NSYS=512, NRHS=4)
    assert t == 0.5 and calls == [1]
    line = path.read_text().strip()
    assert line.startswith("ID=1 ROUT=dposv_batched T=16 WPC=2 N=64 NSYS=512 NRHS=4 TIME=")
    assert line.endswith("MFLOP=200.0")
    sm2 = m.Sum(str(path), "dposv_batched")  # a later run reads it back
    assert sm2.get(lambda: calls.append(2) or 9.0, 1.0, T=16, WPC=2, N=64, NSYS=512,
                   NRHS=4) == 0.5
    assert calls == [1]
