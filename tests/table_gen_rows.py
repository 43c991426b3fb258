"""Rows of the ice lookup-table file: reading the committed table and formatting
generator rows in its e15.5 layout (create_p3_lookupTable_1.f90-v4.1:1374-1427)."""
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TABLE = os.path.join(ROOT, "scream_amd", "data", "p3_lookup_table_1.dat-v4.1.1")
N_RF, N_Q, N_D = 20, 50, 30


def e15_5(x: float) -> str:
    """Fortran e15.5: 0.ddddd E+xx with five correctly rounded digits, width 15."""
    if x == 0:
        return "0.00000E+00".rjust(15)
    m, e = ("%.4e" % abs(x)).split("e")
    e = int(e) + 1
    return (("-" if x < 0 else "") + "0." + m.replace(".", "") + "E" + ("+" if e >= 0 else "-")
            + "%02d" % abs(e)).rjust(15)


def table_lines(path=TABLE):
    with open(path) as fh:
        lines = fh.read().split("\n")
    assert lines[0] == "VERSION 4.1.1"
    return [l for l in lines[1:] if l]


def ice_line(lines, cell):
    rf, q = divmod(cell, N_Q)
    return lines[rf * (N_Q + N_Q * N_D) + q]


def coll_line(lines, cell, d):
    rf, q = divmod(cell, N_Q)
    return lines[rf * (N_Q + N_Q * N_D) + N_Q + q * N_D + d]


def format_ice(cell, row):
    rf = cell // N_Q
    return "%5d%5d" % (rf // 4 + 1, rf % 4 + 1) + "".join(e15_5(v) for v in row)


def format_coll(cell, row):
    return format_ice(cell, row)


def parse(lines):
    """The committed table's values as generator rows (ice [1000, 15], coll [1000, 30, 6])."""
    ice = np.zeros((N_RF * N_Q, 15))
    coll = np.zeros((N_RF * N_Q, N_D, 6))
    for c in range(N_RF * N_Q):
        ice[c] = [float(ice_line(lines, c)[10 + 15 * j: 25 + 15 * j]) for j in range(15)]
        for d in range(N_D):
            l = coll_line(lines, c, d)
            coll[c, d] = [float(l[10 + 15 * j: 25 + 15 * j]) for j in range(6)]
    return ice, coll


def mismatches(ice, coll, cells, lines):
    """Rows of the given cells whose formatting differs from the committed file."""
    bad = []
    for c in cells:
        if format_ice(c, ice[c]) != ice_line(lines, c):
            bad.append(("ice", c, format_ice(c, ice[c]), ice_line(lines, c)))
        for d in range(N_D):
            if format_coll(c, coll[c, d]) != coll_line(lines, c, d):
                bad.append(("coll", c, d, format_coll(c, coll[c, d]), coll_line(lines, c, d)))
    return bad
