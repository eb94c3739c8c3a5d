#!/usr/bin/env python3
"""Writes tests/golden/runs_std_atm.npz from the reference's RRTMG_LW
standard-atmosphere cases (doc_rrtm/runs_std_atm/input_rrtm_<case> and
output_rrtm_<case>, the clear-sky ones).  Data only: the column records of
each input (record 2.1.1 / 2.1.2 of doc_rrtm/rrtmg_lw_instructions.txt: layer
pressure and temperature, interface pressures and temperatures, the seven
molecular amounts and the broadening-gas column, and the surface temperature
of record 1.4) and the broadband flux table of each output (level, pressure,
upward, downward and net flux, heating rate).

    python tests/golden/make_runs_std_atm_golden.py /root/reference
"""
import os
import re
import sys

import numpy as np

CASES = ("MLS-clr", "MLW-clr", "SAW-clr", "TROP-clr")


def parse_input(path):
    L = open(path).read().splitlines()
    i = [k for k, l in enumerate(L) if re.match(r"^ [01]\s*\d+\s+\d+", l)][0]   # record 2.1
    iform, nl, nmol = int(L[i][1]), int(L[i][2:5]), int(L[i][5:10])
    tbound = float(L[i - 1].split()[0])                                         # record 1.4
    w = 15 if iform == 1 else 10
    pave, tave, pz, tz, wkl, wbr = [], [], [], [], [], []
    for n in range(nl):
        a, b = L[i + 1 + 2 * n], L[i + 2 + 2 * n]
        pave.append(float(a[0:w]))
        tave.append(float(a[w:w + 10]))
        o = w + 10 + 23
        if n == 0:
            pz.append(float(a[o:o + 8]))
            tz.append(float(a[o + 8:o + 15]))
        o += 15 + 7
        pz.append(float(a[o:o + 8]))
        tz.append(float(a[o + 8:o + 15]))
        ww = 15 if iform == 1 else 10
        vals = [float(b[k * ww:(k + 1) * ww]) for k in range(nmol + 1)]
        wkl.append(vals[:nmol])
        wbr.append(vals[nmol])
    return dict(tbound=tbound, pave=np.array(pave), tave=np.array(tave), pz=np.array(pz), tz=np.array(tz),
                wkl=np.array(wkl), wbrodl=np.array(wbr))


def parse_output(path):
    L = open(path).read().splitlines()
    i = [k for k, l in enumerate(L) if "Wavenumbers" in l][0]
    rows = []
    for l in L[i + 3:]:
        t = l.split()
        if len(t) != 6:
            break
        rows.append([float(x) for x in t])
    a = np.array(rows)[::-1]          # surface first
    return dict(level=a[:, 0].astype(np.int32), p=a[:, 1], up=a[:, 2], down=a[:, 3], net=a[:, 4], hr=a[:, 5])


def main(ref):
    d = os.path.join(ref, "doc_rrtm", "runs_std_atm")
    out = {}
    for c in CASES:
        for k, v in parse_input(os.path.join(d, "input_rrtm_" + c)).items():
            out["%s/in/%s" % (c, k)] = v
        for k, v in parse_output(os.path.join(d, "output_rrtm_" + c)).items():
            out["%s/out/%s" % (c, k)] = v
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runs_std_atm.npz")
    np.savez_compressed(dst, **out)
    print(dst, len(out))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
