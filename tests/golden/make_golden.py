"""Writes the committed fixtures under tests/golden/.

1. atlas_test_params.json: the parameter sets of ATLAS's own tester runs,
   parsed from the reference's test data (read as text when the reference
   tree is present):
     - AtlasBase/Clint/misc/CBLATST/c_dblat3.dat (netlib CBLAS level-3 tester:
       N values, alpha/beta values, threshold),
     - AtlasTest/makes/binrun (range sweeps 1..195 step 27, llttst -T 16).
2. lcg_words.json: ATL_rand words for a set of seeds, computed here from the
   mathematical definition of the generator (a 31-bit LCG with
   a = 16838*2^16 + 20077, c = 12345, first state one step from
   s = iseed + 2^16*(iseed >> 16); AtlasBase/Clint/atlas-tst.base:5598-5615),
   using Python integers, i.e. independently of the C restatement.

Run:  python tests/golden/make_golden.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
A = 16838 * 65536 + 20077
C = 12345
M31 = 1 << 31


def lcg_words(seed, n):
    s = seed + 65536 * (seed >> 16)          # jseed halves, low half unmasked
    x = (s * A + C) % M31
    out = []
    for _ in range(n):
        out.append(x)
        x = (A * x + C) % M31
    return out


def parse_dblat3(path):
    with open(path) as f:
        lines = [ln.split() for ln in f]
    thresh = float(lines[6][0])
    nn = int(lines[7][0])
    nvals = [int(v) for v in lines[8][:nn]]
    na = int(lines[9][0])
    alphas = [float(v) for v in lines[10][:na]]
    nb = int(lines[11][0])
    betas = [float(v) for v in lines[12][:nb]]
    return {"source": "AtlasBase/Clint/misc/CBLATST/c_dblat3.dat", "threshold": thresh,
            "N": nvals, "alpha": alphas, "beta": betas}


def parse_binrun(path):
    with open(path) as f:
        txt = f.read()
    out = {"source": "AtlasTest/makes/binrun"}
    m = re.search(r"xdl3blastst\S*\s+(.*)", txt)
    if m:
        out["l3blastst_args"] = m.group(1).split(">")[0].strip()
    m = re.search(r"xdlutst\S*\s+(.*)", txt)
    if m:
        out["lutst_args"] = m.group(1).split(">")[0].strip()
    m = re.search(r"xdllttst\S*\s+(.*)", txt)
    if m:
        out["llttst_args"] = m.group(1).split(">")[0].strip()
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    params_path = os.path.join(HERE, "atlas_test_params.json")
    if os.path.isdir(ref):
        params = {
            "cblat3": parse_dblat3(os.path.join(ref, "AtlasBase/Clint/misc/CBLATST/c_dblat3.dat")),
            "binrun": parse_binrun(os.path.join(ref, "AtlasTest/makes/binrun")),
            "thresholds": {"l3blastst": 50.0, "lutst": 100.0, "llttst": 100.0,
                           "llttst_atlastest": 16.0, "slvtst": 100.0},
        }
        with open(params_path, "w") as f:
            json.dump(params, f, indent=1)
    seeds = [0, 1, 7, 1234, 65535, 65536, 70000, 1000 * 1000 + 1000,
             1000 * 1000 + 513 * 7 + 90, 1000 * 1000 + 127 * 50 + 77,
             1000 * 1000 + 101 * 2 + 53, 195 * 1029 + 195, 2147483647, -5]
    words = {str(s): lcg_words(s, 16) for s in seeds}
    with open(os.path.join(HERE, "lcg_words.json"), "w") as f:
        json.dump({"generator": "ATL_rand (atlas-tst.base:5353-5620), closed form", "words": words},
                  f, indent=1)


if __name__ == "__main__":
    main()
