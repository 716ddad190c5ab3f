"""Writes the golden fixtures under tests/golden/ from data files the reference holds (run in the build container,
where the reference is mounted; the fixtures are committed and the GPU box never reads the reference).

* strictmath_sincos.txt: the sin/cos lines of the reference's Java StrictMath vector file
  (tests/GridSim.Tests/fixtures/strictmath/strictmath-vectors.txt): explicit hex vectors and SHA-256 digest lines.
* matpower/case*.m: MATPOWER case files the reference bundles (data/matpower/lib), the real-grid inputs.
* worked_example_case9.json: the numbers printed in the reference's sparse-LU worked example
  (docs/technical/0004 Sparse and GPU numerics.md:778-838) that the oracle must reproduce.
* cli/*.txt: the reference's recorded CLI output of its contingency screens
  (tests/GridSim.Tests/fixtures/cli-stdout: nminus1-case9, nminus1-case14, nminus2-case9).
"""
import json
import os
import re
import shutil
import sys

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    src = os.path.join(REF, "tests/GridSim.Tests/fixtures/strictmath/strictmath-vectors.txt")
    lines = [ln for ln in open(src, encoding="utf-8") if re.match(r"^[vd] (sin|cos) ", ln)]
    with open(os.path.join(HERE, "strictmath_sincos.txt"), "w", encoding="utf-8") as f:
        f.write("# sin/cos lines of the reference's Java StrictMath vector file (FDLIBM 5.3 results).\n")
        f.writelines(lines)
    for name in ("nminus1-case9", "nminus1-case14", "nminus2-case9"):
        shutil.copy(os.path.join(REF, "tests/GridSim.Tests/fixtures/cli-stdout", name + ".txt"),
                    os.path.join(HERE, "cli", name + ".txt"))
    for case in ("case9", "case14", "case118", "case300"):
        shutil.copy(os.path.join(REF, "data/matpower/lib", case + ".m"), os.path.join(HERE, "matpower", case + ".m"))
    doc = open(os.path.join(REF, "docs/technical/0004 Sparse and GPU numerics.md"), encoding="utf-8").read()
    q = re.search(r"RCM ordering q \(pivot step k -> original column\): ([0-9 ]+)", doc).group(1).split()
    p = re.search(r"pivot row p \(pivot step k -> original row\):\s+([0-9 ]+)", doc).group(1).split()
    ukk = re.findall(r"^\s*\d+\s+\d+\s+\d+\s+(-?[0-9.]+E[+-]\d+)\s*$", doc, re.M)[:14]
    nnz = re.search(r"nnz\(L\) incl. unit diagonal = (\d+), nnz\(U\) incl. pivots = (\d+)", doc)
    maxj = re.search(r"max \|J_ij\| = ([0-9.]+)", doc).group(1)
    struct = re.search(r"Jacobian order (\d+), structural nnz (\d+) of \d+ \([0-9.]+ %\), numerically non-zero (\d+)",
                       doc)
    q14 = re.search(r"case14 pass 1 iteration 1.*?RCM ordering q: ([0-9 ]+)", doc, re.S).group(1).split()
    fact14 = re.search(r"nnz\(L\)\+nnz\(U\)-n natural = (\d+), RCM = (\d+)", doc)
    out = {
        "source": "docs/technical/0004 Sparse and GPU numerics.md:778-838",
        "case9": {"order": int(struct.group(1)), "structural_nnz": int(struct.group(2)),
                  "numeric_nnz": int(struct.group(3)), "max_abs_j": float(maxj),
                  "q": [int(v) for v in q], "pivot_row": [int(v) for v in p],
                  "u_diag": [float(v) for v in ukk], "nnz_l": int(nnz.group(1)), "nnz_u": int(nnz.group(2))},
        "case14": {"q_rcm": [int(v) for v in q14], "factor_entries_natural": int(fact14.group(1)),
                   "factor_entries_rcm": int(fact14.group(2))},
    }
    json.dump(out, open(os.path.join(HERE, "worked_example_case9.json"), "w"), indent=1)
    print("fixtures written", file=sys.stderr)


if __name__ == "__main__":
    main()
