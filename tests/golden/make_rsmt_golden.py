"""Writes rsmt_powv.json: random point sets of 4..9 distinct gcells on grids
with random gcell lengths, and the minimal rectilinear Steiner tree length of
each, from FLUTE's wirelength-vector table of the reference tree
(run/POWV9.dat, evaluated by tests/powv.py).  Run from the repository root:
    python tests/golden/make_rsmt_golden.py [/root/reference/run/POWV9.dat]
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
from powv import POWV  # noqa: E402


def cases(seed, per_degree):
    rng = random.Random(seed)
    out = []
    for d in range(4, 10):
        for _ in range(per_degree):
            X, Y = rng.randint(d, 40), rng.randint(d, 40)
            xlen = [rng.randint(1, 9) for _ in range(X - 1)]
            ylen = [rng.randint(1, 9) for _ in range(Y - 1)]
            pts = set()
            # some sets share rows or columns (gaps of zero in FLUTE's terms)
            span = rng.choice([X, max(2, X // 3)])
            while len(pts) < d:
                pts.add((rng.randrange(min(X, span)), rng.randrange(Y)))
            pts = sorted(pts)
            out.append({"xlen": xlen, "ylen": ylen, "x": [p[0] for p in pts], "y": [p[1] for p in pts]})
    return out


def main():
    table = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/run/POWV9.dat"
    p = POWV(table)
    rows = cases(2024, 60)
    for r in rows:
        drx = [0]
        for v in r["xlen"]:
            drx.append(drx[-1] + v)
        dry = [0]
        for v in r["ylen"]:
            dry.append(dry[-1] + v)
        r["opt"] = p.length([drx[x] for x in r["x"]], [dry[y] for y in r["y"]])
    with open(os.path.join(HERE, "rsmt_powv.json"), "w") as f:
        json.dump({"source": "minimal lengths from run/POWV9.dat of the reference tree (tests/powv.py)",
                   "cases": rows}, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
