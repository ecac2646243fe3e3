#!/usr/bin/env python3
"""Extract RRTMG_SW's raw data tables from the reference's Fortran sources into
one binary blob, ``rrtmg_sw_tables.bin``, that the HIP library embeds and the
oracle reads.

Nothing is computed here beyond what the Fortran data statements themselves do:
every ``name(sect) = (/ v1, v2, ... /)`` array constructor and every scalar
``name = v`` assignment is parsed as text and stored at the array position the
Fortran column-major layout gives it. The only arithmetic is the band-29
``irradnceo`` rescale the data file applies to itself
(src/rrtmg_sw_k_g.f90:62460-62461), done in double precision as the reference's
``-r8`` build does (build/makefiles/make_rrtmg_sw_linux_pgi:43).

Sources read (text only):
  * src/rrtmg_sw_k_g.f90        raw 16-g k-tables and solar source per band
                                (sw_kgb16 :53 ... sw_kgb29 :62433)
  * src/rrtmg_sw_init.f90       swaerpr :389-489 (ECMWF aerosol),
                                swcldpr :1692-3514 (cloud optics)
Shapes follow modules/rrsw_kg16.f90 ... rrsw_kg29.f90, rrsw_cld.f90:40-44 and
rrsw_aer.f90.

Blob layout (little endian):
  char magic[8] = "RRTMGSWT"; uint32 version = 1; uint32 nrec;
  nrec x { char name[32]; int32 ndim; int32 lbound[4]; int32 extent[4];
           int64 offset_doubles; }
  float64 data[...]      (each record column-major, Fortran order)

Usage:  python extract_tables.py [/root/reference] [out.bin]
"""
import hashlib
import os
import re
import struct
import sys

import numpy as np

NO = 16  # raw g-points per band

# name -> list of (lbound, ubound) per dimension, exactly as declared in
# modules/rrsw_kgNN.f90 (prefixed kgNN_), rrsw_cld.f90 and rrsw_aer.f90.
def _kg_shapes():
    s = {}
    sp = {
        16: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 rayl=[]),
        17: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 4), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 5)], irradnceo=[(1, NO), (1, 5)],
                 facbrghto=[(1, NO), (1, 5)], snsptdrko=[(1, NO), (1, 5)], rayl=[]),
        18: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 9)], irradnceo=[(1, NO), (1, 9)],
                 facbrghto=[(1, NO), (1, 9)], snsptdrko=[(1, NO), (1, 9)], rayl=[]),
        19: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 9)], irradnceo=[(1, NO), (1, 9)],
                 facbrghto=[(1, NO), (1, 9)], snsptdrko=[(1, NO), (1, 9)], rayl=[]),
        20: dict(kao=[(1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 4), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 absch4o=[(1, NO)], rayl=[]),
        21: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 4), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 9)], irradnceo=[(1, NO), (1, 9)],
                 facbrghto=[(1, NO), (1, 9)], snsptdrko=[(1, NO), (1, 9)], rayl=[]),
        22: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 9)], irradnceo=[(1, NO), (1, 9)],
                 facbrghto=[(1, NO), (1, 9)], snsptdrko=[(1, NO), (1, 9)], rayl=[]),
        23: dict(kao=[(1, 5), (1, 13), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 raylo=[(1, NO)]),
        24: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 3), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 9)], irradnceo=[(1, NO), (1, 9)],
                 facbrghto=[(1, NO), (1, 9)], snsptdrko=[(1, NO), (1, 9)],
                 abso3ao=[(1, NO)], abso3bo=[(1, NO)], raylao=[(1, NO), (1, 9)], raylbo=[(1, NO)]),
        25: dict(kao=[(1, 5), (1, 13), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 abso3ao=[(1, NO)], abso3bo=[(1, NO)], raylo=[(1, NO)]),
        26: dict(sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 raylo=[(1, NO)]),
        27: dict(kao=[(1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 raylo=[(1, NO)]),
        28: dict(kao=[(1, 9), (1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (1, 5), (13, 59), (1, NO)],
                 sfluxrefo=[(1, NO), (1, 5)], irradnceo=[(1, NO), (1, 5)],
                 facbrghto=[(1, NO), (1, 5)], snsptdrko=[(1, NO), (1, 5)], rayl=[]),
        29: dict(kao=[(1, 5), (1, 13), (1, NO)], kbo=[(1, 5), (13, 59), (1, NO)],
                 selfrefo=[(1, 10), (1, NO)], forrefo=[(1, 4), (1, NO)],
                 sfluxrefo=[(1, NO)], irradnceo=[(1, NO)], facbrghto=[(1, NO)], snsptdrko=[(1, NO)],
                 absh2oo=[(1, NO)], absco2o=[(1, NO)], rayl=[]),
    }
    for b, d in sp.items():
        for k, v in d.items():
            s[f"kg{b}_{k}"] = v
    return s


CLD_SHAPES = {
    **{n: [(1, 5)] for n in ("abari", "bbari", "cbari", "dbari", "ebari", "fbari")},
    **{n: [(1, 58), (16, 29)] for n in ("extliq1", "ssaliq1", "asyliq1")},
    **{n: [(1, 43), (16, 29)] for n in ("extice2", "ssaice2", "asyice2")},
    **{n: [(1, 46), (16, 29)] for n in ("extice3", "ssaice3", "asyice3", "fdlice3")},
}
AER_SHAPES = {n: [(1, 14), (1, 6)] for n in ("rsrtaua", "rsrpiza", "rsrasya")}

_NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?(?:_rb)?")


def _parse_number(tok):
    tok = tok.strip()
    if tok.endswith("_rb"):
        tok = tok[:-3]
    return float(tok.replace("d", "e").replace("D", "e"))


def _statements(lines):
    """Join Fortran free-form continuation lines; drop comments."""
    buf = ""
    for raw in lines:
        line = raw.split("!")[0].rstrip()
        if not line.strip():
            continue
        s = line.strip()
        if s.startswith("&"):
            s = s[1:]
        if buf:
            buf += " " + s
        else:
            buf = s
        if buf.endswith("&"):
            buf = buf[:-1]
            continue
        yield buf
        buf = ""
    if buf:
        yield buf


class Array:
    def __init__(self, dims):
        self.dims = dims
        self.ext = [u - l + 1 for l, u in dims]
        n = int(np.prod(self.ext)) if dims else 1
        self.data = np.full(n, np.nan)
        self.hits = np.zeros(n, dtype=np.int32)

    def assign(self, sect, values):
        """sect: list of ints or ':' per dimension (Fortran order)."""
        if not self.dims:
            assert sect is None and len(values) == 1
            self.data[0] = values[0]
            self.hits[0] += 1
            return
        assert len(sect) == len(self.dims), (sect, self.dims)
        free = [i for i, s in enumerate(sect) if s == ":"]
        assert len(free) <= 1
        if free:
            fd = free[0]
            lo, hi = self.dims[fd]
            idxs = range(lo, hi + 1)
        else:
            fd, idxs = None, [None]
        assert len(values) == len(idxs), (sect, len(values), len(idxs))
        for v, i in zip(values, idxs):
            pos = list(sect)
            if fd is not None:
                pos[fd] = i
            lin, stride = 0, 1
            for (l, u), p in zip(self.dims, pos):
                assert l <= p <= u, (pos, self.dims)
                lin += (p - l) * stride
                stride *= u - l + 1
            self.data[lin] = v
            self.hits[lin] += 1


_ASSIGN = re.compile(r"^([a-zA-Z_][a-zA-Z_0-9]*)\s*(\(([^=]*)\))?\s*=\s*(.*)$")


def _parse_section(txt):
    out = []
    for p in txt.split(","):
        p = p.strip()
        out.append(":" if p == ":" else int(p))
    return out


def parse_block(lines, prefix, shapes, arrays):
    """Parse the data statements of one Fortran routine body."""
    irradscl = None
    for st in _statements(lines):
        m = _ASSIGN.match(st)
        if not m:
            continue
        name, _, sect, rhs = m.groups()
        lname = name.lower()
        if lname == "irradscl":
            a, b = re.findall(r"([\d.]+)_rb", rhs)[0], re.findall(r"- *([\d.]+)_rb", rhs)[0]
            irradscl = float(a) / (float(a) - float(b))
            continue
        key = prefix + lname
        if key not in shapes:
            continue
        arr = arrays.setdefault(key, Array(shapes[key]))
        rhs = rhs.strip()
        if rhs.startswith("(/"):
            body = rhs[2:rhs.rindex("/)")]
            vals = [_parse_number(t) for t in body.split(",") if t.strip()]
            arr.assign(_parse_section(sect) if sect is not None else [":"] * len(shapes[key]), vals)
        elif "irradscl" in rhs:
            assert irradscl is not None and lname == "irradnceo"
            arr.data *= irradscl
        else:
            arr.assign(None if not shapes[key] else _parse_section(sect), [_parse_number(rhs)])


def extract(ref_root):
    shapes = _kg_shapes()
    shapes.update(CLD_SHAPES)
    shapes.update(AER_SHAPES)
    arrays = {}
    with open(os.path.join(ref_root, "src", "rrtmg_sw_k_g.f90")) as f:
        kg = f.readlines()
    starts = {}
    for i, ln in enumerate(kg):
        m = re.match(r"\s+subroutine sw_kgb(\d+)", ln)
        if m:
            starts[int(m.group(1))] = i
    for b in range(16, 30):
        i0 = starts[b]
        i1 = next(i for i in range(i0, len(kg)) if re.match(r"\s+end subroutine sw_kgb", kg[i]))
        parse_block(kg[i0:i1], f"kg{b}_", shapes, arrays)
    with open(os.path.join(ref_root, "src", "rrtmg_sw_init.f90")) as f:
        ini = f.readlines()

    def body(name):
        i0 = next(i for i, l in enumerate(ini) if re.match(rf"\s+subroutine {name}\b", l))
        i1 = next(i for i in range(i0, len(ini)) if re.match(rf"\s+end subroutine {name}\b", ini[i]))
        return ini[i0:i1]

    parse_block(body("swcldpr"), "", shapes, arrays)
    parse_block(body("swaerpr"), "", shapes, arrays)
    missing = sorted(set(shapes) - set(arrays))
    if missing:
        raise SystemExit(f"arrays never assigned: {missing}")
    for k, a in arrays.items():
        if not (a.hits == 1).all():
            raise SystemExit(f"{k}: element assigned {a.hits.min()}..{a.hits.max()} times")
        if np.isnan(a.data).any():
            raise SystemExit(f"{k}: NaN")
    return arrays


def write_blob(arrays, path):
    names = sorted(arrays)
    header = bytearray()
    header += b"RRTMGSWT" + struct.pack("<II", 1, len(names))
    off = 0
    recs = []
    for n in names:
        a = arrays[n]
        lb = [l for l, _ in a.dims] + [1] * (4 - len(a.dims))
        ex = a.ext + [1] * (4 - len(a.ext))
        header += struct.pack("<32si4i4iq", n.encode(), len(a.dims), *lb, *ex, off)
        recs.append(a.data)
        off += a.data.size
    with open(path, "wb") as f:
        f.write(bytes(header))
        np.concatenate(recs).astype("<f8").tofile(f)
    return off


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(here, "rrtmg_sw_tables.bin")
    arrays = extract(ref)
    n = write_blob(arrays, out)
    with open(out, "rb") as f:
        digest = hashlib.sha256(f.read()).hexdigest()
    print(f"wrote {out}: {len(arrays)} arrays, {n} doubles, sha256 {digest}")


if __name__ == "__main__":
    main()
