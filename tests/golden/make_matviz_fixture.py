"""Convert the reference's dumped 5-point system (devDocs/matviz/data.txt,
vin.txt, vout.txt in the reference repo) into tests/golden/matviz_64.npz.

Text files only are read (the reference's Python scripts are not run):
data.txt rows are "index:diag iNeg iPos jNeg jPos" on a 64x64 grid, vin/vout
one value per line (solvetest.py:9-80 documents the layout).
"""
import os
import re
import sys

import numpy as np

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/devDocs/matviz"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "matviz_64.npz")

num = r"(-?\d+\.*\d*(?:e-?\d+)?)"
row = re.compile(rf"(\d+):{num} {num} {num} {num} {num}")
rows, size = [], None
for line in open(os.path.join(SRC, "data.txt")):
    if "=" in line:
        continue
    m = row.search(line)
    if m:
        rows.append([float(g) for g in m.groups()])
    elif line.count(",") == 1:
        size = tuple(int(x) for x in line.split(","))


def vec(name):
    return np.array([float(x) for x in open(os.path.join(SRC, name)) if "=" not in x and x.strip()], np.float64)


np.savez_compressed(OUT, rows=np.array(rows, np.float64), size=np.array(size, np.int64), vin=vec("vin.txt"),
                    vout=vec("vout.txt"))
print(f"wrote {OUT}: {len(rows)} rows, grid {size}")
