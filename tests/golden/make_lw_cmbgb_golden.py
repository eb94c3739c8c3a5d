#!/usr/bin/env python3
"""Write the RRTMG_LW g-point reduction literals as a fixture.

Reads the reference's ``rrtmg_lw_init.f90`` as text (nothing is compiled or
run) and stores, in ``lw_cmbgb_literals.npz`` next to this script:

* ``ngc``, ``ngs``, ``ngm``, ``ngn``, ``wt`` -- the data statements of
  ``lwcmbdat`` (``rrtmg_lw_init.f90:284-363``), the reduction's mapping and
  the 16 RRTM quadrature weights;
* per band b = 1..16, from the text of ``cmbgb<b>`` (``:366-2015``):
  ``rwgt_off[b-1]`` -- the constant k in ``rwgt(iprsm+k)`` (band b reads the
  weights of its own 16 original g-points), ``ngn_off[b-1]`` -- the ngs entry
  in ``ngn(ngs(n)+igc)`` (0 for ``ngn(igc)``), and the original arrays the
  routine reduces: ``weighted_<b>`` (accumulated as ``x*rwgt``) and
  ``plain_<b>`` (accumulated without a weight: the Planck fractions).

Run in a checkout that has the reference tree:

    python tests/golden/make_lw_cmbgb_golden.py /root/reference

The parser here is independent of ``mima_amd/tables/extract_tables.py``, so
the CPU tests can check that script's LW bookkeeping against these literals.
"""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
INIT = "src/atmos_param/rrtm_radiation/rrtmg_lw/gcm_model/src/rrtmg_lw_init.f90"


def _code(text):
    """Fortran text with comments removed and continuation lines joined."""
    lines = []
    for raw in text.splitlines():
        s = raw.split("!", 1)[0].rstrip()
        if s.strip():
            lines.append(s)
    joined = "\n".join(lines)
    return re.sub(r"&\s*\n\s*&?", " ", joined)


def _body(code, name):
    m = re.search(r"^\s*subroutine\s+%s\b(.*?)^\s*end\s+subroutine\s+%s\b" % (name, name), code,
                  re.M | re.S | re.I)
    if not m:
        raise KeyError(name)
    return m.group(1)


def _list(code, name):
    m = re.search(r"\b%s\s*\(\s*:\s*\)\s*=\s*\(/(.*?)/\)" % name, code, re.S)
    if not m:
        raise KeyError(name)
    return [t.strip().replace("_rb", "") for t in m.group(1).split(",") if t.strip()]


def main(ref):
    code = _code(open(os.path.join(ref, INIT)).read())
    dat = _body(code, "lwcmbdat")
    out = {
        "ngc": np.array([int(v) for v in _list(dat, "ngc")], np.int32),
        "ngs": np.array([int(v) for v in _list(dat, "ngs")], np.int32),
        "ngm": np.array([int(v) for v in _list(dat, "ngm")], np.int32),
        "ngn": np.array([int(v) for v in _list(dat, "ngn")], np.int32),
        "wt": np.array([float(v) for v in _list(dat, "wt")], np.float64),
    }
    assert out["ngc"].size == 16 and out["ngm"].size == 256 and out["ngn"].size == 140 and out["wt"].size == 16
    rwgt_off, ngn_off = [], []
    for b in range(1, 17):
        body = _body(code, "cmbgb%d" % b)
        offs = set(int(k or 0) for k in re.findall(r"rwgt\(\s*iprsm\s*(?:\+\s*(\d+))?\s*\)", body))
        if len(offs) != 1:
            raise ValueError("cmbgb%d: rwgt offsets %s" % (b, offs))
        rwgt_off.append(offs.pop())
        nofs = set(int(k or 0) for k in re.findall(r"ngn\(\s*(?:ngs\(\s*(\d+)\s*\)\s*\+\s*)?igc\s*\)", body))
        if len(nofs) != 1:
            raise ValueError("cmbgb%d: ngn offsets %s" % (b, nofs))
        ngn_off.append(nofs.pop())
        weighted = sorted(set(re.findall(r"\+\s*(\w+)\s*\([^()]*\)\s*\*\s*rwgt\(", body)))
        plain = sorted(set(re.findall(r"=\s*sumf\w*\s*\+\s*(\w+)\s*\([^()]*\)\s*$", body, re.M)))
        if not weighted or not plain:
            raise ValueError("cmbgb%d: no reductions found" % b)
        out["weighted_%d" % b] = np.array(weighted)
        out["plain_%d" % b] = np.array(plain)
    out["rwgt_off"] = np.array(rwgt_off, np.int32)
    out["ngn_off"] = np.array(ngn_off, np.int32)
    path = os.path.join(HERE, "lw_cmbgb_literals.npz")
    np.savez(path, **out)
    print("wrote", path)
    for b in range(1, 17):
        print(b, rwgt_off[b - 1], ngn_off[b - 1], list(out["weighted_%d" % b]), list(out["plain_%d" % b]))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
