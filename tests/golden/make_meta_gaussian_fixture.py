"""Writes tests/golden/meta_gaussian_coefs.json from the reference's own
coefficient generator (tools/make_meta_gaussian/make_meta_gaussian.cpp), built
from its source by `make -C oracle ref` into oracle/_ref/.  Only the output data
(sigma, x, the printed coefficient literal) is stored."""
import json
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
OUT = os.path.join(ROOT, "tests", "golden", "meta_gaussian_coefs.json")
PAT = re.compile(r"meta_kernel<meta_gaussian<(\d+), (\d+)>, (-?\d+)>.*?return ([^;]+);")


def parse(path, order=0):
    rows = []
    with open(path) as f:
        for line in f:
            m = PAT.search(line)
            if m and int(m.group(1)) == order:
                rows.append((int(m.group(2)), int(m.group(3)), m.group(4).strip()))
    return rows


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    data = {}
    for sigma in (1, 2, 3, 4, 5, 8):
        hdr = os.path.join(ROOT, "oracle", "_ref", "meta_gaussian", f"meta_gaussian_coefs_{sigma}.h")
        rows = parse(hdr)
        K = 3 * sigma + 3
        data[str(sigma)] = {str(x): lit for s, x, lit in rows if abs(x) <= K}
    with open(OUT, "w") as f:
        json.dump({"source": "tools/make_meta_gaussian/make_meta_gaussian.cpp -s (0th derivative)",
                   "coefs": data}, f, indent=1, sort_keys=True)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
