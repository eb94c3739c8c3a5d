#!/usr/bin/env python3
"""Build the RRTMG table files this package loads at init.

Run once in a checkout that has the MiMA reference tree; the outputs are
committed, so nothing downstream needs the Fortran sources.

    python mima_amd/tables/extract_tables.py --ref /root/reference

What it writes (all little-endian, FP64 values exactly as the Fortran
compiler would parse each ``_rb`` literal, i.e. correctly rounded doubles):

* ``rrtmg_sw.tbl`` -- AER RRTMG_SW data: the original 16-g-point k-tables of
  bands 16-29 (``rrtmg_sw_k_g.f90``), the band/g-point bookkeeping of
  ``swdatinit``/``swcmbdat`` (``rrtmg_sw_init.f90:157-367``) and the MLS
  reference atmosphere of ``swatmref`` (``rrtmg_sw_setcoef.f90:289-343``).
* ``rrtmg_lw_cld.tbl`` / ``rrtmg_sw_cld.tbl`` -- the cloud optics
  coefficients of ``lwcldpr`` (``rrtmg_lw_init.f90``) and ``swcldpr``
  (``rrtmg_sw_init.f90``), for cldprop's water-path cloud optics.
* ``rrtmg_lw.tbl`` -- RRTMG_LW bookkeeping (``rrtmg_lw_init.f90:178-363``),
  reference atmosphere and ``chi_mls`` (``rrtmg_lw_setcoef.f90:418-578``),
  the Planck table ``totplnk`` (``rrtmg_lw_setcoef.f90:581-1284``), the
  diffusivity fit ``a0/a1/a2`` (``rrtmg_lw_rtrnmr.f90:246-262``) and
  **synthetic** 16-g-point k-tables.  AER's LW k-coefficient file
  (``rrtmg_lw_k_g.f90``) is not available, so every LW absorption
  coefficient here comes from a seeded generator that only reproduces the
  array shapes declared in ``modules/rrlw_kg01..16.f90``.  Entries are named
  ``lw/bNN/...`` and the file header carries ``synthetic_kg=1``.

The g-point reduction (16 -> ngc per band) is *not* done here; the library
and the oracle each do it at load time from these original-resolution arrays.

Format ``RRTMGTB1``::

    char[8]  magic "RRTMGTB1"
    uint32   n_entries
    uint32   flags              (bit 0: LW k-tables are synthetic)
    entry[n_entries]:
        char[48] name           (NUL padded, e.g. "sw/b16/kao")
        uint32   dtype          (0 = float64, 1 = int32)
        uint32   ndim
        int32    lbound[4]      (Fortran lower bounds, unused dims = 1)
        int32    extent[4]      (unused dims = 1; first index fastest)
        uint64   offset         (bytes from start of file)
        uint64   nbytes
    payload (each blob 64-byte aligned)

The RRTMG data are Copyright (c) 2002-2009 Atmospheric and Environmental
Research, Inc. (AER).  They are redistributed here with that notice and
without charge, as the AER licence in the RRTMG source headers requires.
"""
import argparse
import hashlib
import json
import os
import re
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
RR = "src/atmos_param/rrtm_radiation"
SW = RR + "/rrtmg_sw/gcm_model"
LW = RR + "/rrtmg_lw/gcm_model"

MAGIC = b"RRTMGTB1"
HDR = struct.Struct("<8sII")
ENT = struct.Struct("<48sII4i4iQQ")

# ----------------------------------------------------------------------------
# Fortran source scanning
# ----------------------------------------------------------------------------

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?(?:_rb|_im)?")


def _statements(text):
    """Yield free-form Fortran statements with comments and '&' joins removed."""
    buf = ""
    for raw in text.splitlines():
        line = raw.split("!", 1)[0].rstrip()
        s = line.strip()
        if not s:
            continue
        if s.startswith("&"):
            s = s[1:]
        if s.endswith("&"):
            buf += s[:-1] + " "
            continue
        buf += s
        yield buf
        buf = ""
    if buf:
        yield buf


def _subroutine_body(text, name):
    m = re.search(r"^\s*subroutine\s+%s\b.*?$(.*?)^\s*end\s+subroutine\s+%s\b" % (name, name),
                  text, re.M | re.S | re.I)
    if not m:
        raise KeyError("subroutine %s not found" % name)
    return m.group(1)


def _parse_number(tok):
    tok = tok.replace("_rb", "").replace("_im", "").replace("d", "e").replace("D", "e")
    if re.fullmatch(r"[-+]?\d+", tok):
        return int(tok)
    return float(tok)


def _parse_list(body):
    # body is the text between (/ and /)
    toks = [t.strip() for t in body.split(",")]
    return [_parse_number(t) for t in toks if t]


class FortranArrays:
    """Collects array-section assignments into numpy arrays with Fortran bounds."""

    def __init__(self):
        self.arrays = {}      # name -> (np.ndarray (Fortran order), lbounds)

    def declare(self, name, shape_spec, dtype=np.float64):
        lb, ext = [], []
        for d in shape_spec:
            if isinstance(d, tuple):
                lb.append(d[0])
                ext.append(d[1] - d[0] + 1)
            else:
                lb.append(1)
                ext.append(d)
        self.arrays[name] = (np.full(ext, np.nan if dtype == np.float64 else -999999,
                                     dtype=dtype, order="F"), lb)

    def assign(self, name, subs, values):
        arr, lb = self.arrays[name]
        if subs is None:
            subs = [":"] * arr.ndim
        idx = []
        for d, s in enumerate(subs):
            s = s.strip()
            if s == ":":
                idx.append(slice(0, arr.shape[d]))
            elif ":" in s:
                a, b = s.split(":")
                idx.append(slice(int(a) - lb[d], int(b) - lb[d] + 1))
            else:
                idx.append(int(s) - lb[d])
        view = arr[tuple(idx)]
        vals = np.asarray(values, dtype=arr.dtype)
        if view.size != vals.size:
            raise ValueError("%s%s: %d slots, %d values" % (name, subs, view.size, vals.size))
        # Fortran fills an array section in column-major order
        arr[tuple(idx)] = vals.reshape(view.shape, order="F")

    def scan(self, text, names):
        """Apply every `name(...) = (/ ... /)` and `name = number` for the given names."""
        pat = re.compile(r"^\s*(\w+)\s*(\(([^=]*)\))?\s*=\s*(.*)$", re.S)
        hits = 0
        for st in _statements(text):
            m = pat.match(st)
            if not m or m.group(1) not in names:
                continue
            name, subs, rhs = m.group(1), m.group(3), m.group(4).strip()
            if rhs.startswith("(/"):
                body = rhs[2:rhs.rindex("/)")]
                vals = _parse_list(body)
            else:
                vals = [_parse_number(rhs)]
            self.assign(name, subs.split(",") if subs else None, vals)
            hits += 1
        return hits


def _module_decls(text, params):
    """Parse `real(kind=rb) :: a(5,13:59,no16), b(10)` style declarations."""
    decls = {}
    for st in _statements(text):
        if "::" not in st or not re.match(r"\s*(real|integer)", st, re.I):
            continue
        if "parameter" in st.lower():
            continue
        rhs = st.split("::", 1)[1]
        dimm = re.search(r"dimension\s*\(([^)]*)\)", st, re.I)
        for m in re.finditer(r"(\w+)\s*(\(([^)]*)\))?", rhs):
            name, dims = m.group(1), m.group(3)
            if dims is None and dimm is not None:
                dims = dimm.group(1)
            spec = []
            if dims:
                for d in dims.split(","):
                    d = d.strip()
                    if ":" in d:
                        a, b = d.split(":")
                        spec.append((int(params.get(a, a)), int(params.get(b, b))))
                    else:
                        spec.append(int(params.get(d, d)))
            decls[name] = spec
    return decls


def _module_params(text):
    return {m.group(1): int(m.group(2)) for m in
            re.finditer(r"parameter\s*::\s*(\w+)\s*=\s*(\d+)", text, re.I)}


def _read(ref, rel):
    with open(os.path.join(ref, rel)) as f:
        return f.read()


def _sha(ref, rel):
    with open(os.path.join(ref, rel), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


# ----------------------------------------------------------------------------
# SW
# ----------------------------------------------------------------------------

SW_BANDS = list(range(16, 30))


def extract_sw(ref):
    out = {}
    provenance = {}
    init = _read(ref, SW + "/src/rrtmg_sw_init.f90")
    provenance[SW + "/src/rrtmg_sw_init.f90"] = _sha(ref, SW + "/src/rrtmg_sw_init.f90")

    fa = FortranArrays()
    for n in ("wavenum1", "wavenum2", "delwave"):
        fa.declare(n, [14])
    for n in ("nspa", "nspb", "ng", "ngc", "ngs"):
        fa.declare(n, [14], np.int32)
    fa.declare("ngm", [14 * 16], np.int32)
    fa.declare("ngn", [112], np.int32)
    fa.declare("ngb", [112], np.int32)
    fa.declare("wt", [16])
    fa.scan(_subroutine_body(init, "swdatinit"), {"wavenum1", "wavenum2", "delwave", "nspa", "nspb", "ng"})
    fa.scan(_subroutine_body(init, "swcmbdat"), {"ngc", "ngs", "ngm", "ngn", "ngb", "wt"})
    for n in ("wavenum1", "wavenum2", "delwave", "nspa", "nspb", "ngc", "ngs", "ngm", "ngn", "ngb", "wt"):
        out["sw/" + n] = fa.arrays[n][0]

    setc = _read(ref, SW + "/src/rrtmg_sw_setcoef.f90")
    provenance[SW + "/src/rrtmg_sw_setcoef.f90"] = _sha(ref, SW + "/src/rrtmg_sw_setcoef.f90")
    fr = FortranArrays()
    for n in ("pref", "preflog", "tref"):
        fr.declare(n, [59])
    fr.scan(_subroutine_body(setc, "swatmref"), {"pref", "preflog", "tref"})
    for n in ("pref", "preflog", "tref"):
        out["sw/" + n] = fr.arrays[n][0]

    kg = _read(ref, SW + "/src/rrtmg_sw_k_g.f90")
    provenance[SW + "/src/rrtmg_sw_k_g.f90"] = _sha(ref, SW + "/src/rrtmg_sw_k_g.f90")
    for b in SW_BANDS:
        modrel = SW + "/modules/rrsw_kg%d.f90" % b
        mod = _read(ref, modrel)
        provenance[modrel] = _sha(ref, modrel)
        params = dict(_module_params(_read(ref, SW + "/modules/parrrsw.f90")), **_module_params(mod))
        decls = _module_decls(mod, params)
        body = _subroutine_body(kg, "sw_kgb%d" % b)
        # the arrays sw_kgbNN fills are exactly those it imports from rrsw_kgNN
        use = re.search(r"use\s+rrsw_kg%d\s*,\s*only\s*:(.*?)(?:\n\s*\n|use\s)" % b,
                        body.replace("&\n", " "), re.S | re.I)
        onames = [t.strip() for t in use.group(1).replace("&", " ").split(",") if t.strip()]
        fb = FortranArrays()
        for n in onames:
            fb.declare(n, decls[n] if decls[n] else [1])
        fb.scan(body, set(onames))
        for n in onames:
            arr = fb.arrays[n][0]
            if np.isnan(arr).any():
                raise ValueError("sw band %d: %s has unset entries" % (b, n))
            out["sw/b%d/%s" % (b, n)] = (arr, fb.arrays[n][1])
    return out, provenance


# ----------------------------------------------------------------------------
# LW
# ----------------------------------------------------------------------------

LW_BANDS = list(range(1, 17))


def extract_lw_data(ref):
    out = {}
    provenance = {}
    init = _read(ref, LW + "/src/rrtmg_lw_init.f90")
    provenance[LW + "/src/rrtmg_lw_init.f90"] = _sha(ref, LW + "/src/rrtmg_lw_init.f90")
    fa = FortranArrays()
    for n in ("wavenum1", "wavenum2", "delwave"):
        fa.declare(n, [16])
    for n in ("nspa", "nspb", "ng", "ngc", "ngs"):
        fa.declare(n, [16], np.int32)
    fa.declare("ngm", [16 * 16], np.int32)
    fa.declare("ngn", [140], np.int32)
    fa.declare("ngb", [140], np.int32)
    fa.declare("wt", [16])
    fa.scan(_subroutine_body(init, "lwdatinit"), {"wavenum1", "wavenum2", "delwave", "nspa", "nspb", "ng"})
    fa.scan(_subroutine_body(init, "lwcmbdat"), {"ngc", "ngs", "ngm", "ngn", "ngb", "wt"})
    for n in ("wavenum1", "wavenum2", "delwave", "nspa", "nspb", "ngc", "ngs", "ngm", "ngn", "ngb", "wt"):
        out["lw/" + n] = fa.arrays[n][0]

    setc = _read(ref, LW + "/src/rrtmg_lw_setcoef.f90")
    provenance[LW + "/src/rrtmg_lw_setcoef.f90"] = _sha(ref, LW + "/src/rrtmg_lw_setcoef.f90")
    fr = FortranArrays()
    for n in ("pref", "preflog", "tref"):
        fr.declare(n, [59])
    fr.declare("chi_mls", [7, 59])
    fr.scan(_subroutine_body(setc, "lwatmref"), {"pref", "preflog", "tref", "chi_mls"})
    for n in ("pref", "preflog", "tref", "chi_mls"):
        if np.isnan(fr.arrays[n][0]).any():
            raise ValueError("lw %s has unset entries" % n)
        out["lw/" + n] = fr.arrays[n][0]
    fp = FortranArrays()
    fp.declare("totplnk", [181, 16])
    fp.scan(_subroutine_body(setc, "lwavplank"), {"totplnk"})
    if np.isnan(fp.arrays["totplnk"][0]).any():
        raise ValueError("totplnk has unset entries")
    out["lw/totplnk"] = fp.arrays["totplnk"][0]

    rt = _read(ref, LW + "/src/rrtmg_lw_rtrnmr.f90")
    provenance[LW + "/src/rrtmg_lw_rtrnmr.f90"] = _sha(ref, LW + "/src/rrtmg_lw_rtrnmr.f90")
    for n in ("a0", "a1", "a2"):
        m = re.search(r"data\s+%s\s*/(.*?)/" % n, rt.replace("&\n", " "), re.S)
        vals = [_parse_number(t.strip()) for t in m.group(1).replace("&", " ").split(",") if t.strip()]
        if len(vals) != 16:
            raise ValueError("%s: %d values" % (n, len(vals)))
        out["lw/" + n] = np.asarray(vals, dtype=np.float64)
    return out, provenance


def extract_lw_cloud(ref):
    """The LW cloud absorption coefficients of lwcldpr (rrtmg_lw_init.f90,
    module rrlw_cld): abscld1, absice0(2), absice1(2,5), absice2(43,16),
    absice3(46,16), absliq0, absliq1(58,16), for cldprop's inflglw = 1, 2."""
    init = _read(ref, LW + "/src/rrtmg_lw_init.f90")
    prov = {LW + "/src/rrtmg_lw_init.f90": _sha(ref, LW + "/src/rrtmg_lw_init.f90")}
    fa = FortranArrays()
    shapes = {"abscld1": [1], "absice0": [2], "absice1": [2, 5], "absice2": [43, 16], "absice3": [46, 16],
              "absliq0": [1], "absliq1": [58, 16]}
    for n, sh in shapes.items():
        fa.declare(n, sh)
    fa.scan(_subroutine_body(init, "lwcldpr"), set(shapes))
    out = {}
    for n in shapes:
        a = fa.arrays[n][0]
        if np.isnan(a).any():
            raise ValueError("lwcld %s has unset entries" % n)
        out["lwcld/" + n] = a
    return out, prov


def extract_sw_cloud(ref):
    """The SW cloud optics coefficients of swcldpr (rrtmg_sw_init.f90,
    module rrsw_cld) for cldprop_sw's inflgsw = 2, bands 16-29 as columns
    1-14, and the bands' upper wavenumbers (swdatinit) that pick the
    iceflgsw = 1 coefficients."""
    init = _read(ref, SW + "/src/rrtmg_sw_init.f90")
    prov = {SW + "/src/rrtmg_sw_init.f90": _sha(ref, SW + "/src/rrtmg_sw_init.f90")}
    fa = FortranArrays()
    shapes = {}
    for n in ("extliq1", "ssaliq1", "asyliq1"):
        shapes[n] = [58, (16, 29)]
    for n in ("extice2", "ssaice2", "asyice2"):
        shapes[n] = [43, (16, 29)]
    for n in ("extice3", "ssaice3", "asyice3", "fdlice3"):
        shapes[n] = [46, (16, 29)]
    for n in ("abari", "bbari", "cbari", "dbari", "ebari", "fbari"):
        shapes[n] = [5]
    for n, sh in shapes.items():
        fa.declare(n, sh)
    fa.scan(_subroutine_body(init, "swcldpr"), set(shapes))
    fw = FortranArrays()
    fw.declare("wavenum2", [(16, 29)])
    fw.scan(_subroutine_body(init, "swdatinit"), {"wavenum2"})
    out = {}
    for n in shapes:
        a = fa.arrays[n][0]
        if np.isnan(a).any():
            raise ValueError("swcld %s has unset entries" % n)
        out["swcld/" + n] = a
    if np.isnan(fw.arrays["wavenum2"][0]).any():
        raise ValueError("sw wavenum2 has unset entries")
    out["swcld/wavenum2"] = fw.arrays["wavenum2"][0]
    return out, prov


def lw_shapes(ref):
    """Declared shapes of the original-resolution (16 g) LW arrays per band."""
    shapes = {}
    provenance = {}
    for b in LW_BANDS:
        modrel = LW + "/modules/rrlw_kg%02d.f90" % b
        mod = _read(ref, modrel)
        provenance[modrel] = _sha(ref, modrel)
        params = dict(_module_params(_read(ref, LW + "/modules/parrrtm.f90")), **_module_params(mod))
        decls = _module_decls(mod, params)
        # original-resolution arrays: kao, kbo, kao_m*, kbo_m* and the other
        # names ending in "o"; reduced minor tables ka_m*/kb_m* may also end in "o"
        shapes[b] = {n: s for n, s in decls.items() if s and (
            n.startswith(("kao", "kbo")) or (n.endswith("o") and not n.startswith(("ka_", "kb_"))))}
    return shapes, provenance


def synth_lw_kg(shapes, seed):
    """Seeded, shape-exact stand-in values for the missing AER LW k-tables.

    Magnitudes are log-uniform per g-point (so optical depths span thin to
    opaque), smooth in pressure/temperature/eta, and Planck fractions are
    positive and sum to one over the 16 original g-points of each (band, eta).
    These are NOT physical coefficients.
    """
    rng = np.random.default_rng(seed)
    out = {}
    for b in LW_BANDS:
        for name, spec in sorted(shapes[b].items()):
            lb, ext = [], []
            for d in spec:
                if isinstance(d, tuple):
                    lb.append(d[0]); ext.append(d[1] - d[0] + 1)
                else:
                    lb.append(1); ext.append(d)
            if name.startswith("fracref"):
                # fracrefao(no, neta): the g-point index is the first dimension
                w = rng.uniform(0.5, 1.5, size=ext)
                arr = np.asfortranarray(w / w.sum(axis=0, keepdims=True))
            else:
                ng = ext[-1]
                g = np.arange(ng, dtype=np.float64)
                # base magnitude rises with g (low g = window, high g = line centre)
                if name.startswith(("kao", "kbo")) and not name.startswith(("kao_", "kbo_")):
                    lo, hi = -4.0, 2.5
                elif name.startswith(("selfref", "forref")):
                    lo, hi = -5.0, -1.5
                elif name.startswith(("kao_", "kbo_")):
                    lo, hi = -6.0, -1.0
                else:   # cross sections (cfc*, ccl4)
                    lo, hi = -2.0, 1.0
                base = lo + (hi - lo) * (g / max(ng - 1, 1))
                shape_other = ext[:-1]
                smooth = np.zeros(ext)
                for ax, n in enumerate(shape_other):
                    coord = np.linspace(-1.0, 1.0, n) if n > 1 else np.zeros(1)
                    amp = rng.uniform(0.1, 0.6)
                    phase = rng.uniform(0, np.pi)
                    prof = amp * np.sin(1.3 * coord + phase)
                    sh = [1] * len(ext)
                    sh[ax] = n
                    smooth = smooth + prof.reshape(sh)
                jitter = rng.normal(0.0, 0.05, size=ext)
                arr = np.asfortranarray(10.0 ** (base.reshape([1] * len(shape_other) + [ng]) + smooth + jitter))
            out["lw/b%02d/%s" % (b, name)] = (arr, lb)
    return out


# ----------------------------------------------------------------------------
# Writer / reader
# ----------------------------------------------------------------------------

def write_table(path, entries, flags):
    names = sorted(entries)
    blobs = []
    for n in names:
        v = entries[n]
        arr, lb = (v if isinstance(v, tuple) else (v, None))
        arr = np.asarray(arr)
        if arr.dtype.kind == "f":
            dtype, data = 0, np.asarray(arr, dtype="<f8")
        else:
            dtype, data = 1, np.asarray(arr, dtype="<i4")
        ext = list(arr.shape)
        if lb is None:
            lb = [1] * len(ext)
        if len(ext) > 4:
            raise ValueError(n)
        blobs.append((n, dtype, len(ext), lb + [1] * (4 - len(lb)), ext + [1] * (4 - len(ext)),
                      data.tobytes(order="F")))
    off = HDR.size + ENT.size * len(blobs)
    off = (off + 63) // 64 * 64
    header = [HDR.pack(MAGIC, len(blobs), flags)]
    payload = []
    cur = off
    for n, dt, nd, lb, ext, raw in blobs:
        header.append(ENT.pack(n.encode().ljust(48, b"\0"), dt, nd, *lb, *ext, cur, len(raw)))
        pad = (-len(raw)) % 64
        payload.append(raw + b"\0" * pad)
        cur += len(raw) + pad
    head = b"".join(header)
    head += b"\0" * (off - len(head))
    with open(path, "wb") as f:
        f.write(head)
        for p in payload:
            f.write(p)


def read_table(path):
    """Return ({name: np.ndarray (Fortran order)}, {name: lbounds}, flags)."""
    with open(path, "rb") as f:
        raw = f.read()
    magic, n, flags = HDR.unpack_from(raw, 0)
    if magic != MAGIC:
        raise ValueError("%s: not an RRTMGTB1 file" % path)
    arrays, bounds = {}, {}
    for i in range(n):
        name, dt, nd, *rest = ENT.unpack_from(raw, HDR.size + i * ENT.size)
        lb, ext, off, nb = rest[0:4], rest[4:8], rest[8], rest[9]
        name = name.rstrip(b"\0").decode()
        a = np.frombuffer(raw, dtype="<f8" if dt == 0 else "<i4", count=nb // (8 if dt == 0 else 4),
                          offset=off)
        arrays[name] = a.reshape(ext[:nd], order="F")
        bounds[name] = list(lb[:nd])
    return arrays, bounds, flags


def write_shapes_header(path, table_dir=HERE):
    """C++ header listing the shape of every band array of the two table
    files (the reference's module declarations, as the files hold them): the
    loader refuses a table file whose arrays differ, since the band kernels
    index them with those extents."""
    lines = ["// Generated by mima_amd/tables/extract_tables.py --shapes-header from the",
             "// committed table files: name, rank and extents (first index fastest) of",
             "// every band array, as declared by the reference's rrlw_kgNN / rrsw_kgNN",
             "// modules.  The table loader refuses a file whose arrays differ.",
             "#pragma once", "", "namespace rrtmg {", "",
             "struct TableShape {", "    const char *name;", "    int ndim, ext[4];", "};", ""]
    for side in ("sw", "lw"):
        arrays, _, _ = read_table(os.path.join(table_dir, "rrtmg_%s.tbl" % side))
        names = sorted(n for n in arrays if n.count("/") == 2)
        lines.append("constexpr TableShape %s_SHAPES[] = {" % side.upper())
        for n in names:
            ext = list(arrays[n].shape) + [1] * (4 - arrays[n].ndim)
            lines.append('    {"%s", %d, {%s}},' % (n, arrays[n].ndim, ", ".join(str(e) for e in ext)))
        lines.append("};")
        lines.append("constexpr int %s_NSHAPES = (int)(sizeof(%s_SHAPES) / sizeof(%s_SHAPES[0]));"
                     % ((side.upper(),) * 3))
        lines.append("")
    lines.append("}  // namespace rrtmg")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--out", default=HERE)
    ap.add_argument("--lw-seed", type=int, default=20261202)
    ap.add_argument("--lw-cloud-only", action="store_true",
                    help="only write rrtmg_lw_cld.tbl and rrtmg_sw_cld.tbl (lwcldpr's and swcldpr's cloud "
                         "optics coefficients)")
    ap.add_argument("--shapes-header", default=None,
                    help="only write the band-array shape header (C++) of the committed tables to this path")
    args = ap.parse_args()
    if args.shapes_header:
        write_shapes_header(args.shapes_header)
        return 0
    if args.lw_cloud_only:
        mpath = os.path.join(args.out, "manifest.json")
        manifest = json.load(open(mpath)) if os.path.exists(mpath) else {}
        for name, fn in (("rrtmg_lw_cld.tbl", extract_lw_cloud), ("rrtmg_sw_cld.tbl", extract_sw_cloud)):
            cld, prov = fn(args.ref)
            path = os.path.join(args.out, name)
            write_table(path, cld, 0)
            manifest[name] = {"sha256": hashlib.sha256(open(path, "rb").read()).hexdigest(), "sources": prov}
            print("wrote", len(cld), "entries to", name)
        with open(mpath, "w") as f:
            json.dump(manifest, f, indent=1, sort_keys=True)
        return 0

    sw, prov_sw = extract_sw(args.ref)
    write_table(os.path.join(args.out, "rrtmg_sw.tbl"), sw, 0)

    cld, prov_cld = extract_lw_cloud(args.ref)
    write_table(os.path.join(args.out, "rrtmg_lw_cld.tbl"), cld, 0)
    scld, prov_scld = extract_sw_cloud(args.ref)
    write_table(os.path.join(args.out, "rrtmg_sw_cld.tbl"), scld, 0)

    lw, prov_lw = extract_lw_data(args.ref)
    shapes, prov_sh = lw_shapes(args.ref)
    lw.update(synth_lw_kg(shapes, args.lw_seed))
    write_table(os.path.join(args.out, "rrtmg_lw.tbl"), lw, 1)

    manifest = {
        "format": "RRTMGTB1",
        "notice": "RRTMG data Copyright (c) 2002-2009 Atmospheric and Environmental Research, Inc. (AER). "
                  "Redistributed with this notice and without charge.",
        "rrtmg_sw.tbl": {"sha256": hashlib.sha256(open(os.path.join(args.out, "rrtmg_sw.tbl"), "rb").read()).hexdigest(),
                          "sources": prov_sw},
        "rrtmg_lw.tbl": {"sha256": hashlib.sha256(open(os.path.join(args.out, "rrtmg_lw.tbl"), "rb").read()).hexdigest(),
                          "sources": dict(prov_lw, **prov_sh),
                          "synthetic_kg": True, "synthetic_seed": args.lw_seed,
                          "synthetic_note": "AER rrtmg_lw_k_g.f90 is absent; LW k-tables are seeded "
                                            "shape-exact synthetic values, not physics."},
        "rrtmg_lw_cld.tbl": {"sha256": hashlib.sha256(open(os.path.join(args.out, "rrtmg_lw_cld.tbl"),
                                                           "rb").read()).hexdigest(), "sources": prov_cld},
        "rrtmg_sw_cld.tbl": {"sha256": hashlib.sha256(open(os.path.join(args.out, "rrtmg_sw_cld.tbl"),
                                                           "rb").read()).hexdigest(), "sources": prov_scld},
    }
    with open(os.path.join(args.out, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    write_shapes_header(os.path.join(HERE, "..", "csrc", "table_shapes.hpp"), args.out)
    print("wrote", len(sw), "SW entries and", len(lw), "LW entries")


if __name__ == "__main__":
    sys.exit(main())
