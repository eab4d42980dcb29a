#!/usr/bin/env python3
"""Regenerate ``scoef.dat`` from the SRIM coefficient tables shipped with MyTRIM.

The engine needs, per element Z = 1..92, the columns that
``SimconfType::readDataFiles`` parses from ``data/SCOEF.95A`` and
``data/SLFCTR.dat`` (reference: simconf.C:82-137, simconf.C:182-215) and that
the stopping code reads (material.C:221-224, 251-253):

    mm1 m1 mnat rho atrho vfermi heat pcoef[0..7] lfctr

The source files are whitespace token streams with two (SCOEF.95A) or one
(SLFCTR.dat) header lines, read exactly like ``operator>>`` does in the
reference.  Values are re-printed with 17 significant digits so that every
double round-trips bit-exactly.  SCOEF.95B and SNUC03.dat are parsed by the
reference but never used on any code path, so they are not carried.

Usage: python gen_tables.py /root/reference/data  > scoef.dat
"""
import os
import sys


def read_tables(datadir):
    with open(os.path.join(datadir, "SCOEF.95A")) as f:
        lines = f.read().split("\n")
    toks = " ".join(lines[2:]).split()
    rows = []
    pos = 0
    for z in range(1, 93):
        zz = int(toks[pos])
        if zz != z:
            raise ValueError("SCOEF.95A row %d has Z=%d" % (z, zz))
        vals = [float(t) for t in toks[pos + 1:pos + 16]]
        pos += 16
        rows.append(vals)  # mm1 m1 mnat rho atrho vfermi heat pcoef*8
    with open(os.path.join(datadir, "SLFCTR.dat")) as f:
        lines = f.read().split("\n")
    toks = " ".join(lines[1:]).split()
    for z in range(1, 93):
        zz = int(toks[2 * (z - 1)])
        if zz != z:
            raise ValueError("SLFCTR.dat row %d has Z=%d" % (z, zz))
        rows[z - 1].append(float(toks[2 * (z - 1) + 1]))
    return rows


def main():
    datadir = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/data"
    rows = read_tables(datadir)
    out = sys.stdout
    out.write("# Z mm1 m1 mnat rho atrho vfermi heat pcoef0..pcoef7 lfctr\n")
    for z, r in enumerate(rows, start=1):
        out.write("%d %s\n" % (z, " ".join("%.17g" % v for v in r)))


if __name__ == "__main__":
    main()
