"""Optimal rectilinear Steiner tree lengths for 4..9 points from FLUTE's
potentially-optimal wirelength vectors (POWV9.dat), as a test checker.

FLUTE's table lists, for every degree d <= 9 and every group of point
permutations, the wirelength vectors that can be optimal: the tree length is
the half perimeter plus a sum of gaps between consecutive sorted coordinates,
and the minimum over the group's vectors is the exact minimal length (Chu and
Wong, TCAD 2008).  This module restates, in Python:
  - the table parser: readLUTfiles (src/flute.hpp:298-380), POWV part only;
  - the coordinate sort of flute_wl (src/flute.hpp:711-810, duplicate-pin
    removal off, FLUTE_REMOVE_DUPLICATE_PIN 0 at :66);
  - the group index and vector minimum of flutes_wl_LD (src/flute.hpp:861-930).
It reads the table from the reference tree; only the fixture generator and
the CPU tests (when the reference tree is present) use it.
"""
NUMGRP = [0, 0, 0, 0, 6, 30, 180, 1260, 10080, 90720]   # src/flute.hpp:255


def _charnum(c):
    if ord("0") <= c <= ord("9"):
        return c - ord("0")
    if c >= ord("A"):
        return c - ord("A") + 10
    return 0


class POWV:
    def __init__(self, path, maxd=9):
        data = open(path, "rb").read()
        pos = 0
        self.lut = {}

        def getc():
            nonlocal pos
            c = data[pos]
            pos += 1
            return c

        def read_int():
            nonlocal pos
            while data[pos] in b" \t\r\n":
                pos += 1
            st = pos
            while ord("0") <= data[pos] <= ord("9"):
                pos += 1
            return int(data[st:pos])

        for d in range(4, maxd + 1):
            assert data[pos:pos + 2] == b"d=", "table out of step at degree %d" % d
            pos += 2
            assert read_int() == d
            getc()   # newline
            groups = []
            for k in range(NUMGRP[d]):
                ns = _charnum(getc())
                if ns == 0:
                    kk = read_int()
                    getc()
                    groups.append(groups[kk])
                    continue
                getc()
                sols = []
                for _ in range(ns):
                    end = data.index(b"\n", pos)
                    line = data[pos:end + 1]
                    pos = end + 1
                    parent = _charnum(line[0])
                    add, i = [], 1
                    while True:
                        v = _charnum(line[i])
                        i += 1
                        if v == 0:
                            break
                        add.append(v)
                    sub = []
                    while i < len(line):
                        v = _charnum(line[i])
                        i += 1
                        if v == 0:
                            break
                        sub.append(v)
                    sols.append((parent, add, sub))
                groups.append(sols)
            self.lut[d] = groups

    def length(self, x, y):
        """Minimal rectilinear Steiner tree length of the points (x[i], y[i])."""
        d = len(x)
        if d == 2:
            return abs(x[0] - x[1]) + abs(y[0] - y[1])
        if d == 3:
            return max(x) - min(x) + max(y) - min(y)
        # flute_wl: selection sort by x (first minimum, swapped in), then by y
        pts = [[x[i], y[i], 0] for i in range(d)]
        ptp = list(range(d))
        for i in range(d - 1):
            mi = i
            for j in range(i + 1, d):
                if pts[ptp[mi]][0] > pts[ptp[j]][0]:
                    mi = j
            ptp[i], ptp[mi] = ptp[mi], ptp[i]
        xs = [pts[ptp[i]][0] for i in range(d)]
        for i in range(d):
            pts[ptp[i]][2] = i
        ys, s = [0] * d, [0] * d
        for i in range(d - 1):
            mi = i
            for j in range(i + 1, d):
                if pts[ptp[mi]][1] > pts[ptp[j]][1]:
                    mi = j
            ys[i] = pts[ptp[mi]][1]
            s[i] = pts[ptp[mi]][2]
            ptp[mi] = ptp[i]
        ys[d - 1] = pts[ptp[d - 1]][1]
        s[d - 1] = pts[ptp[d - 1]][2]
        # flutes_wl_LD: group index from the permutation, gaps, vector minimum
        k = (1 if s[0] < s[2] else 0) + (1 if s[1] < s[2] else 0)
        for i in range(3, d):
            pi = s[i]
            for j in range(d - 1, i, -1):
                if s[j] < s[i]:
                    pi -= 1
            k = pi + (i + 1) * k
        dd = [0] * (2 * d)
        if k < NUMGRP[d]:
            for i in range(1, d - 2):
                dd[i] = ys[i + 1] - ys[i]
                dd[d - 1 + i] = xs[i + 1] - xs[i]
        else:
            k = 2 * NUMGRP[d] - 1 - k
            for i in range(1, d - 2):
                dd[i] = ys[i + 1] - ys[i]
                dd[d - 1 + i] = xs[d - 1 - i] - xs[d - 2 - i]
        sols = self.lut[d][k]
        base = xs[d - 1] - xs[0] + ys[d - 1] - ys[0]
        ls = [base]
        minl = base + sum(dd[v] for v in sols[0][1])
        ls.append(minl)
        for parent, add, sub in sols[1:]:
            v = ls[parent] + sum(dd[a] for a in add) - sum(dd[b] for b in sub)
            minl = min(minl, v)
            ls.append(v)
        return minl
