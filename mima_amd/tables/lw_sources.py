#!/usr/bin/env python3
"""AER's RRTMG_LW k-coefficient data in its two published forms, turned into
the table file the library loads (``rrtmg_lw.tbl``, format RRTMGTB1 of
extract_tables.py).

The reference fills the original 16-g-point LW arrays of every band
(``kao``, ``kbo``, ``selfrefo``, ``forrefo``, ``fracrefao``, ``fracrefbo``,
the minor-gas tables ``kao_m*``/``kbo_m*`` and the cross sections) from one
of two sources, both called from ``rrtmg_lw_ini``
(``rrtmg_lw/gcm_model/src/rrtmg_lw_init.f90:80-95``):

* ``rrtmg_lw_k_g.f90``: subroutines ``lw_kgb01..16``, each assigning the
  arrays it imports from module ``rrlw_kgNN`` with array-section literals
  ``kao(:, 1, 1) = (/ 1.2e-04_rb, ... /)`` (the form of the SW file
  ``rrtmg_sw/gcm_model/src/rrtmg_sw_k_g.f90``);
* ``rrtmg_lw.nc``: one netCDF file read hyperslab by hyperslab
  (``rrtmg_lw/gcm_model/src/rrtmg_lw_read_nc.f90:19-1032``) from seven
  variables whose dimensions are given by ``modules/rrlw_ncpar.f90``.

Neither file is in the MiMA tree (``.MISSING_LARGE_BLOBS:3-4``); the
committed ``rrtmg_lw.tbl`` carries seeded stand-in k-tables.  With AER's
data, ``python -m mima_amd.tables.lw_sources --kg rrtmg_lw_k_g.f90 --out DIR``
(or ``--nc rrtmg_lw.nc``) writes ``DIR/rrtmg_lw.tbl`` with the real tables
and every other entry (bookkeeping, Planck table, reference atmosphere) from
the committed file; ``rrtmg_hip_lw_init(cpdair, DIR)`` then loads it.

The netCDF reader handles the classic formats (CDF-1 and the 64-bit-offset
CDF-2) in plain Python; an HDF5-based netCDF-4 file is refused with a
message (no HDF5 library is used here).  Writers for both forms exist so the
loaders can be checked by round trips on the committed tables.
"""
import argparse
import importlib.util
import os
import re
import struct
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _extract_tables():
    spec = importlib.util.spec_from_file_location("extract_tables", os.path.join(HERE, "extract_tables.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


ET = _extract_tables()

# ----------------------------------------------------------------------------
# The schema of rrtmg_lw.nc (rrtmg_lw_read_nc.f90, modules/rrlw_ncpar.f90).
# Dimensions in Fortran order (first fastest), as nf90_get_var's start/count.

NC_DIMS = dict(keylower=9, keyupper=5, Tdiff=5, plower=13, pupper=47, Tself=10, Tforeign=4, T=19, GPoint=16,
               Absorber=12, Bands=16, GPointSet=2)
ABSORBERS = ("N2", "CCL4", "CFC11", "CFC12", "CFC22", "H2O", "CO2", "O3", "N2O", "CO", "CH4", "O2")
NC_VARS = {
    "PlanckFractionLowerAtmos": ("GPoint", "keylower", "Bands", "GPointSet"),
    "PlanckFractionUpperAtmos": ("GPoint", "keyupper", "Bands", "GPointSet"),
    "KeySpeciesAbsorptionCoefficientsLowerAtmos": ("keylower", "Tdiff", "plower", "GPoint", "Bands", "GPointSet"),
    "KeySpeciesAbsorptionCoefficientsUpperAtmos": ("keyupper", "Tdiff", "pupper", "GPoint", "Bands", "GPointSet"),
    "H20SelfAbsorptionCoefficients": ("Tself", "GPoint", "Bands", "GPointSet"),
    "H20ForeignAbsorptionCoefficients": ("Tforeign", "GPoint", "Bands", "GPointSet"),
    "AbsorptionCoefficientsLowerAtmos": ("keylower", "T", "GPoint", "Absorber", "Bands", "GPointSet"),
    "AbsorptionCoefficientsUpperAtmos": ("keyupper", "T", "GPoint", "Absorber", "Bands", "GPointSet"),
}
_PF = {"fracrefao": "PlanckFractionLowerAtmos", "fracrefbo": "PlanckFractionUpperAtmos"}
_KEY = {"kao": "KeySpeciesAbsorptionCoefficientsLowerAtmos", "kbo": "KeySpeciesAbsorptionCoefficientsUpperAtmos"}
# (array, variable, absorber or None, count of the leading dimensions) per band,
# as lw_kgb01..16 of rrtmg_lw_read_nc.f90 read them; the trailing dimensions
# (band, g-point set 1, absorber) are pinned by start
_S, _F = ("selfrefo", "H20SelfAbsorptionCoefficients", None, ("Tself", "GPoint")), \
    ("forrefo", "H20ForeignAbsorptionCoefficients", None, ("Tforeign", "GPoint"))


def _pf(name, key):
    return (name, _PF[name], None, ("GPoint", key) if key else ("GPoint", 1))


def _k(name, key):
    low = name == "kao"
    return (name, _KEY[name], None, (key or 1, "Tdiff", "plower" if low else "pupper", "GPoint"))


def _minor(name, ab, key=None, cross=False):
    low = name.startswith("kao") or cross
    var = "AbsorptionCoefficientsLowerAtmos" if low else "AbsorptionCoefficientsUpperAtmos"
    return (name, var, ab, (1, 1, "GPoint") if cross else (key or 1, "T", "GPoint"))


NC_SCHEMA = {
    1: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F,
        _minor("kao_mn2", "N2"), _minor("kbo_mn2", "N2")],
    2: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F],
    3: [_pf("fracrefao", "keylower"), _pf("fracrefbo", "keyupper"), _k("kao", "keylower"), _k("kbo", "keyupper"),
        _S, _F, _minor("kao_mn2o", "N2O", "keylower"), _minor("kbo_mn2o", "N2O", "keyupper")],
    4: [_pf("fracrefao", "keylower"), _pf("fracrefbo", "keyupper"), _k("kao", "keylower"), _k("kbo", "keyupper"),
        _S, _F],
    5: [_pf("fracrefao", "keylower"), _pf("fracrefbo", "keyupper"), _k("kao", "keylower"), _k("kbo", "keyupper"),
        _S, _F, _minor("kao_mo3", "O3", "keylower"), _minor("ccl4o", "CCL4", cross=True)],
    6: [_pf("fracrefao", None), _k("kao", None), _S, _F, _minor("kao_mco2", "CO2"),
        _minor("cfc11adjo", "CFC11", cross=True), _minor("cfc12o", "CFC12", cross=True)],
    7: [_pf("fracrefao", "keylower"), _pf("fracrefbo", None), _k("kao", "keylower"), _k("kbo", None), _S, _F,
        _minor("kao_mco2", "CO2", "keylower"), _minor("kbo_mco2", "CO2")],
    8: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F,
        _minor("kao_mo3", "O3"), _minor("kao_mco2", "CO2"), _minor("kbo_mco2", "CO2"), _minor("kao_mn2o", "N2O"),
        _minor("kbo_mn2o", "N2O"), _minor("cfc12o", "CFC12", cross=True), _minor("cfc22adjo", "CFC22", cross=True)],
    9: [_pf("fracrefao", "keylower"), _pf("fracrefbo", None), _k("kao", "keylower"), _k("kbo", None), _S, _F,
        _minor("kao_mn2o", "N2O", "keylower"), _minor("kbo_mn2o", "N2O")],
    10: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F],
    11: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F,
         _minor("kao_mo2", "O2"), _minor("kbo_mo2", "O2")],
    12: [_pf("fracrefao", "keylower"), _k("kao", "keylower"), _S, _F],
    13: [_pf("fracrefao", "keylower"), _pf("fracrefbo", None), _k("kao", "keylower"), _S, _F,
         _minor("kbo_mo3", "O3"), _minor("kao_mco2", "CO2", "keylower"), _minor("kao_mco", "CO", "keylower")],
    14: [_pf("fracrefao", None), _pf("fracrefbo", None), _k("kao", None), _k("kbo", None), _S, _F],
    15: [_pf("fracrefao", "keylower"), _k("kao", "keylower"), _S, _F, _minor("kao_mn2", "N2", "keylower")],
    16: [_pf("fracrefao", "keylower"), _pf("fracrefbo", None), _k("kao", "keylower"), _k("kbo", None), _S, _F],
}


def _count(spec):
    return [NC_DIMS[d] if isinstance(d, str) else d for d in spec]


def _hyperslab(var, band, ab, lead):
    """Fortran-order (start, count) of one array's hyperslab of a variable."""
    dims = NC_VARS[var]
    start, count = [], []
    for i, d in enumerate(dims):
        if i < len(lead):
            start.append(1)
            count.append(lead[i])
        elif d == "Absorber":
            start.append(ABSORBERS.index(ab) + 1)
            count.append(1)
        elif d == "Bands":
            start.append(band)
            count.append(1)
        elif d == "GPointSet":
            start.append(1)
            count.append(1)
        else:
            start.append(1)
            count.append(1)
    return start, count


# ----------------------------------------------------------------------------
# The arrays each band needs: names and declared shapes (with Fortran lower
# bounds) from the committed table file, whose loader checks them against the
# reference's module declarations.

def band_arrays(table_path=None):
    """{band: {name: (extent, lbound)}} of the original-resolution LW arrays."""
    arrs, bounds, _ = ET.read_table(table_path or os.path.join(HERE, "rrtmg_lw.tbl"))
    out = {}
    for key, a in arrs.items():
        m = re.fullmatch(r"lw/b(\d\d)/(\w+)", key)
        if not m:
            continue
        out.setdefault(int(m.group(1)), {})[m.group(2)] = (tuple(np.asarray(a).shape), tuple(bounds[key]))
    return out


# ----------------------------------------------------------------------------
# rrtmg_lw_k_g.f90 (array-section literals)

def read_kg_f90(path, table_path=None):
    """{"lw/bNN/name": (array, lbound)} from an rrtmg_lw_k_g.f90-format file.
    Each lw_kgbNN must set every element of every array it imports from
    rrlw_kgNN; anything else raises ValueError."""
    with open(path) as f:
        text = f.read()
    shapes = band_arrays(table_path)
    out = {}
    for b in range(1, 17):
        body = ET._subroutine_body(text, "lw_kgb%02d" % b)
        use = re.search(r"use\s+rrlw_kg%02d\s*,\s*only\s*:(.*?)(?:\n\s*\n|\buse\s|\bimplicit\s)" % b,
                        body.replace("&\n", " "), re.S | re.I)
        if not use:
            raise ValueError("lw_kgb%02d: no 'use rrlw_kg%02d, only:' statement" % (b, b))
        names = [t.strip() for t in use.group(1).replace("&", " ").split(",") if t.strip()]
        names = [n for n in names if n in shapes[b]]
        missing = set(shapes[b]) - set(names)
        if missing:
            raise ValueError("lw_kgb%02d does not import %s" % (b, sorted(missing)))
        fa = ET.FortranArrays()
        for n in names:
            ext, lb = shapes[b][n]
            fa.declare(n, [(l, l + e - 1) for e, l in zip(ext, lb)])
        fa.scan(body, set(names))
        for n in names:
            a = fa.arrays[n][0]
            if np.isnan(a).any():
                raise ValueError("lw_kgb%02d: %s has %d unset entries" % (b, n, int(np.isnan(a).sum())))
            out["lw/b%02d/%s" % (b, n)] = (a, list(fa.arrays[n][1]))
    return out


def _fmt(v):
    s = repr(float(v))
    if "e" not in s and "." not in s:
        s += ".0"
    return s + "_rb"


def write_kg_f90(path, entries):
    """Write entries {"lw/bNN/name": array or (array, lbound)} in the
    rrtmg_lw_k_g.f90 form: one array-section literal per trailing index
    combination, values in shortest round-trip form."""
    lines = ["! RRTMG_LW original-resolution k-coefficients (lw_kgb01..16), written by",
             "! mima_amd/tables/lw_sources.py"]
    for b in range(1, 17):
        names = sorted(k.split("/")[2] for k in entries if k.startswith("lw/b%02d/" % b))
        lines += ["      subroutine lw_kgb%02d" % b, "      use parkind, only : im => kind_im, rb => kind_rb",
                  "      use rrlw_kg%02d, only : %s" % (b, ", ".join(names)), "", "      implicit none", "      save", ""]
        for n in names:
            v = entries["lw/b%02d/%s" % (b, n)]
            a, lb = v if isinstance(v, tuple) else (v, [1] * np.ndim(v))
            a = np.asarray(a, dtype=np.float64)
            if a.ndim == 0:
                a = a.reshape(1)
            tails = np.ndindex(*a.shape[1:]) if a.ndim > 1 else [()]
            for t in tails:
                col = a[(slice(None),) + tuple(t)]
                subs = ", ".join([":"] + [str(i + l) for i, l in zip(t, lb[1:])])
                vals = [_fmt(x) for x in col]
                body = ", &\n        & ".join(", ".join(vals[i:i + 4]) for i in range(0, len(vals), 4))
                lines.append("      %s(%s) = (/ &\n        & %s /)" % (n, subs, body))
        lines += ["", "      end subroutine lw_kgb%02d" % b, ""]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


# ----------------------------------------------------------------------------
# netCDF classic (CDF-1, CDF-2) reader and writer

_NC_TYPES = {1: ("b", 1), 2: ("c", 1), 3: (">i2", 2), 4: (">i4", 4), 5: (">f4", 4), 6: (">f8", 8)}
_NC_DIMENSION, _NC_VARIABLE, _NC_ATTRIBUTE = 10, 11, 12


class _Reader:
    def __init__(self, buf):
        self.b, self.p = buf, 0

    def u32(self):
        v = struct.unpack_from(">I", self.b, self.p)[0]
        self.p += 4
        return v

    def u64(self):
        v = struct.unpack_from(">Q", self.b, self.p)[0]
        self.p += 8
        return v

    def name(self):
        n = self.u32()
        s = self.b[self.p:self.p + n].decode("utf-8")
        self.p += (n + 3) // 4 * 4
        return s

    def attrs(self):
        tag, n = self.u32(), self.u32()
        if tag == 0 and n == 0:
            return {}
        if tag != _NC_ATTRIBUTE:
            raise ValueError("netCDF: bad attribute list")
        out = {}
        for _ in range(n):
            nm = self.name()
            t, cnt = self.u32(), self.u32()
            dt, sz = _NC_TYPES[t]
            raw = self.b[self.p:self.p + cnt * sz]
            self.p += (cnt * sz + 3) // 4 * 4
            out[nm] = raw.decode("utf-8", "replace") if t == 2 else np.frombuffer(raw, dtype=dt)
        return out


def read_netcdf(path):
    """{name: numpy array in C order (netCDF's dimension order)} of every
    variable of a classic-format netCDF file."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:4] == b"\x89HDF":
        raise ValueError("%s is a netCDF-4 (HDF5) file; convert it to the classic format "
                         "(nccopy -k classic) -- this reader handles CDF-1 and CDF-2 only" % path)
    if buf[:3] != b"CDF" or buf[3] not in (1, 2):
        raise ValueError("%s is not a classic netCDF file" % path)
    r = _Reader(buf)
    r.p = 4
    ver = buf[3]
    numrecs = r.u32()
    tag, ndims = r.u32(), r.u32()
    dims = []
    if not (tag == 0 and ndims == 0):
        if tag != _NC_DIMENSION:
            raise ValueError("netCDF: bad dimension list")
        for _ in range(ndims):
            dims.append((r.name(), r.u32()))
    r.attrs()
    tag, nvars = r.u32(), r.u32()
    if not (tag == 0 and nvars == 0) and tag != _NC_VARIABLE:
        raise ValueError("netCDF: bad variable list")
    meta = []
    for _ in range(nvars if tag else 0):
        nm = r.name()
        nd = r.u32()
        ids = [r.u32() for _ in range(nd)]
        r.attrs()
        t, vsize = r.u32(), r.u32()
        begin = r.u32() if ver == 1 else r.u64()
        meta.append((nm, ids, t, vsize, begin))
    recvars = [m for m in meta if m[1] and dims[m[1][0]][1] == 0]
    recsize = sum(m[3] for m in recvars) if len(recvars) > 1 else (recvars[0][3] if recvars else 0)
    out = {}
    for nm, ids, t, vsize, begin in meta:
        dt, sz = _NC_TYPES[t]
        shape = [dims[i][1] for i in ids]
        if ids and dims[ids[0]][1] == 0:       # record variable
            shape[0] = numrecs
            inner = int(np.prod(shape[1:])) if len(shape) > 1 else 1
            rows = [np.frombuffer(buf, dtype=dt, count=inner, offset=begin + k * recsize) for k in range(numrecs)]
            a = np.stack(rows).reshape(shape) if rows else np.zeros(shape)
        else:
            n = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(buf, dtype=dt, count=n, offset=begin).reshape(shape)
        out[nm] = a.astype(np.float64) if t in (3, 4, 5, 6) else a
    return out


def write_netcdf(path, dims, variables, version=2):
    """Classic netCDF (CDF-2 by default) with fixed-size double variables.
    dims: [(name, length)]; variables: [(name, [dim names], C-order array)]."""
    def name(s):
        e = s.encode("utf-8")
        return struct.pack(">I", len(e)) + e + b"\0" * ((4 - len(e) % 4) % 4)
    dix = {d: i for i, (d, _) in enumerate(dims)}
    head = b"CDF" + bytes([version]) + struct.pack(">I", 0)
    head += struct.pack(">II", _NC_DIMENSION, len(dims)) + b"".join(name(d) + struct.pack(">I", n) for d, n in dims)
    head += struct.pack(">II", 0, 0)
    vhdr_len = 8
    for nm, vd, _ in variables:
        vhdr_len += len(name(nm)) + 4 + 4 * len(vd) + 8 + 8 + (4 if version == 1 else 8)
    offset = len(head) + vhdr_len
    vhdr, blobs = struct.pack(">II", _NC_VARIABLE, len(variables)), []
    for nm, vd, a in variables:
        data = np.ascontiguousarray(a, dtype=">f8").tobytes()
        vsize = (len(data) + 3) // 4 * 4
        vhdr += name(nm) + struct.pack(">I", len(vd)) + b"".join(struct.pack(">I", dix[d]) for d in vd)
        vhdr += struct.pack(">II", 0, 0) + struct.pack(">II", 6, vsize)
        vhdr += struct.pack(">I", offset) if version == 1 else struct.pack(">Q", offset)
        blobs.append(data + b"\0" * (vsize - len(data)))
        offset += vsize
    with open(path, "wb") as f:
        f.write(head + vhdr + b"".join(blobs))


def read_kg_nc(path, table_path=None):
    """{"lw/bNN/name": (array, lbound)} from an rrtmg_lw.nc-schema file."""
    v = read_netcdf(path)
    shapes = band_arrays(table_path)
    out = {}
    for b in range(1, 17):
        got = set()
        for name, var, ab, lead in NC_SCHEMA[b]:
            if var not in v:
                raise ValueError("%s: variable %s missing" % (path, var))
            c_arr = v[var]
            want = [NC_DIMS[d] for d in NC_VARS[var]]
            if list(c_arr.shape) != want[::-1]:
                raise ValueError("%s: %s has shape %s, the schema %s" % (path, var, c_arr.shape, want[::-1]))
            start, count = _hyperslab(var, b, ab, _count(lead))
            sl = tuple(slice(s - 1, s - 1 + c) for s, c in zip(start[::-1], count[::-1]))
            flat = np.ascontiguousarray(c_arr[sl]).ravel(order="C")      # = Fortran order of `count`
            ext, lb = shapes[b][name]
            dest = np.full(ext, np.nan, order="F")
            # band 4 reads fracrefbo(:, 1:5): the leading section of the declared array
            n = flat.size
            if n > dest.size:
                raise ValueError("band %d %s: %d values for %d slots" % (b, name, n, dest.size))
            dest.ravel(order="F")[:n] = flat
            dest = np.asfortranarray(flat.reshape(ext, order="F")) if n == dest.size else dest
            if np.isnan(dest).any():
                raise ValueError("band %d %s: the file leaves entries unset" % (b, name))
            out["lw/b%02d/%s" % (b, name)] = (dest, list(lb))
            got.add(name)
        missing = set(shapes[b]) - got
        if missing:
            raise ValueError("band %d: the schema does not supply %s" % (b, sorted(missing)))
    return out


def write_kg_nc(path, entries, version=2):
    """The rrtmg_lw.nc schema filled from entries (g-point set 1; the rest of
    each variable zero)."""
    full = {var: np.zeros([NC_DIMS[d] for d in dims], order="F") for var, dims in NC_VARS.items()}
    for b in range(1, 17):
        for name, var, ab, lead in NC_SCHEMA[b]:
            a = entries["lw/b%02d/%s" % (b, name)]
            a = np.asarray(a[0] if isinstance(a, tuple) else a, dtype=np.float64)
            start, count = _hyperslab(var, b, ab, _count(lead))
            sl = tuple(slice(s - 1, s - 1 + c) for s, c in zip(start, count))
            flat = a.ravel(order="F")[:int(np.prod(count))]
            full[var][sl] = flat.reshape(count, order="F")
    dims = [(d, NC_DIMS[d]) for d in NC_DIMS]
    variables = [(var, list(NC_VARS[var])[::-1], np.ascontiguousarray(full[var].T)) for var in NC_VARS]
    write_netcdf(path, dims, variables, version=version)


# ----------------------------------------------------------------------------
# table file

def build_table(kg, out_dir, base=None):
    """Write out_dir/rrtmg_lw.tbl: the committed table's bookkeeping, Planck
    and reference-atmosphere entries with the k-tables kg; the synthetic flag
    cleared.  Returns the path."""
    base = base or os.path.join(HERE, "rrtmg_lw.tbl")
    arrs, bounds, flags = ET.read_table(base)
    entries = {n: (np.asarray(a), list(bounds[n])) for n, a in arrs.items()}
    for k in [k for k in entries if re.fullmatch(r"lw/b\d\d/\w+", k)]:
        if k not in kg:
            raise ValueError("no data for %s" % k)
    for k, (a, lb) in kg.items():
        old = entries.get(k)
        if old is None or np.asarray(old[0]).shape != np.asarray(a).shape or list(old[1]) != list(lb):
            raise ValueError("%s: shape or bounds differ from the declared array" % k)
        entries[k] = (np.asfortranarray(a), list(lb))
    os.makedirs(out_dir, exist_ok=True)
    path = os.path.join(out_dir, "rrtmg_lw.tbl")
    ET.write_table(path, entries, flags & ~1)
    return path


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    g = ap.add_mutually_exclusive_group(required=True)
    g.add_argument("--kg", help="an rrtmg_lw_k_g.f90-format file")
    g.add_argument("--nc", help="an rrtmg_lw.nc-schema classic netCDF file")
    ap.add_argument("--out", required=True, help="directory for rrtmg_lw.tbl")
    a = ap.parse_args(argv)
    kg = read_kg_f90(a.kg) if a.kg else read_kg_nc(a.nc)
    print(build_table(kg, a.out))


if __name__ == "__main__":
    sys.exit(main())
