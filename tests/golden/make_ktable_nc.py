"""Write a k-table file in the layout src/rrtmg_sw_read_nc.f90 reads (rrtmg_sw.nc), holding the
band data compiled into the library (rrtmg_sw_amd/data/rrtmg_sw_tables.bin, the values of
src/rrtmg_sw_k_g.f90), with scipy's netCDF writer.

    python tests/golden/make_ktable_nc.py out.nc [--version 2] [--extras]

The reference ships no rrtmg_sw.nc, so the file is built here: every value that read_nc takes is
placed at the start/count of its nf90_get_var (netcdf-fortran's start/count are the Fortran order
of the dimensions, the file's own order reversed), and every value it does not take -- other
keys, g-point set 2, other absorbers -- is a seeded junk value far from any table entry, so a
reader that takes one element off its place builds different tables. Dimension names and sizes
follow modules/rrsw_ncpar.f90; the leading two dimensions of the minor-absorber variables are not
named by the shortwave reader (it reads only their first element) and are given here as
(T, keylower).
"""
import argparse
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BLOB = os.path.join(HERE, "..", "..", "rrtmg_sw_amd", "data", "rrtmg_sw_tables.bin")

# modules/rrsw_ncpar.f90
DIMS = dict(keylower=9, keyupper=5, Tdiff=5, plower=13, pupper=47, Tself=10, Tforeignlower=3,
            Tforeignupper=2, T=19, band=14, GPoint=16, GPointSet=2, Absorber=12)
ABSORBERS = ["N2", "CCL4", "CFC11", "CFC12", "CFC22", "H2O", "CO2", "O3", "N2O", "CO", "CH4", "O2"]

# File variables, dimensions in Fortran order (first fastest).
G4 = ("GPoint", "band", "GPointSet")
VARS = {
    "KeySpeciesAbsorptionCoefficientsLowerAtmos": ("keylower", "Tdiff", "plower") + G4,
    "KeySpeciesAbsorptionCoefficientsUpperAtmos": ("keyupper", "Tdiff", "pupper") + G4,
    "H2OSelfAbsorptionCoefficients": ("Tself",) + G4,
    "H2OForeignAbsorptionCoefficientsLowerAtmos": ("Tforeignlower",) + G4,
    "H2OForeignAbsorptionCoefficientsUpperAtmos": ("Tforeignupper",) + G4,
    "AbsorptionCoefficientsLowerAtmos": ("T", "keylower", "GPoint", "Absorber", "band",
                                         "GPointSet"),
    "AbsorptionCoefficientsUpperAtmos": ("T", "keylower", "GPoint", "Absorber", "band",
                                         "GPointSet"),
}
for half, key in (("Lower", "keylower"), ("Upper", "keyupper")):
    for v in ("SolarSourceFunction", "NRLSSI2SSFQuietSun", "NRLSSI2SSFFacular",
              "NRLSSI2SSFSunspot", "RayleighExtinctionCoefficients"):
        VARS[f"{v}{half}Atmos"] = ("GPoint", key, "band", "GPointSet")


def _reads():
    """(band, module array, file variable, Fortran start (1-based), Fortran count, rows) of every
    nf90_get_var in sw_kgb16 .. sw_kgb29 (src/rrtmg_sw_read_nc.f90:46-1047); `rows` is the
    forrefo row section a read fills (None: the whole array)."""
    key_lo, key_up = "KeySpeciesAbsorptionCoefficientsLowerAtmos", \
        "KeySpeciesAbsorptionCoefficientsUpperAtmos"
    out = []

    def add(band, arr, var, start, count, rows=None):
        out.append((band, arr, var, start, count, rows))

    solar_var = {16: "L", 17: "U", 18: "L", 19: "L", 20: "L", 21: "L", 22: "L", 23: "L", 24: "L",
                 25: "L", 26: "L", 27: "L", 28: "U", 29: "L"}
    nrlssi_var = {16: "U", 17: "U", 18: "L", 19: "L", 20: "L", 21: "L", 22: "L", 23: "L",
                  24: "L", 25: "L", 26: "L", 27: "U", 28: "U", 29: "U"}
    solar_keys = {16: 1, 17: 5, 18: 9, 19: 9, 20: 1, 21: 9, 22: 9, 23: 1, 24: 9, 25: 1, 26: 1,
                  27: 1, 28: 5, 29: 1}
    half = {"L": "Lower", "U": "Upper"}
    for b in range(16, 30):
        nk = solar_keys[b]
        add(b, "sfluxrefo", f"SolarSourceFunction{half[solar_var[b]]}Atmos", (1, 1), (16, nk))
        for arr, v in (("irradnceo", "QuietSun"), ("facbrghto", "Facular"),
                       ("snsptdrko", "Sunspot")):
            add(b, arr, f"NRLSSI2SSF{v}{half[nrlssi_var[b]]}Atmos", (1, 1), (16, nk))
    ray = "RayleighExtinctionCoefficients%sAtmos"
    for b in (16, 17, 18, 19, 20, 21, 22, 28, 29):
        add(b, "rayl", ray % "Lower", (1, 1), (1, 1))
    for b in (23, 25, 26, 27):
        add(b, "raylo", ray % "Lower", (1, 1), (16, 1))
    add(24, "raylao", ray % "Lower", (1, 1), (16, 9))
    add(24, "raylbo", ray % "Upper", (1, 1), (16, 1))
    kao_keys = {16: 9, 17: 9, 18: 9, 19: 9, 20: 1, 21: 9, 22: 9, 23: 1, 24: 9, 25: 1, 27: 1,
                28: 9, 29: 1}
    for b, nk in kao_keys.items():
        add(b, "kao", key_lo, (1, 1, 1, 1), (nk, 5, 13, 16))
    kbo_keys = {16: 1, 17: 5, 18: 1, 19: 1, 20: 1, 21: 5, 22: 1, 24: 1, 27: 1, 28: 5, 29: 1}
    for b, nk in kbo_keys.items():
        add(b, "kbo", key_up, (1, 1, 1, 1), (nk, 5, 47, 16))
    for b in (16, 17, 18, 19, 20, 21, 22, 23, 24, 29):
        add(b, "selfrefo", "H2OSelfAbsorptionCoefficients", (1, 1), (10, 16))
    for b in (16, 18, 19, 22, 23, 24):
        add(b, "forrefo", "H2OForeignAbsorptionCoefficientsLowerAtmos", (1, 1), (3, 16))
    for b in (17, 20, 21, 29):   # forrefo(1:3,:) from the lower table, forrefo(4:4,:) upper(2)
        add(b, "forrefo", "H2OForeignAbsorptionCoefficientsLowerAtmos", (1, 1), (3, 16), (0, 3))
        add(b, "forrefo", "H2OForeignAbsorptionCoefficientsUpperAtmos", (2, 1), (1, 16), (3, 4))
    ab = {n: i + 1 for i, n in enumerate(ABSORBERS)}      # getAbsorberIndex
    for b, arr, half_, name in ((20, "absch4o", "Lower", "CH4"), (24, "abso3ao", "Lower", "O3"),
                                (24, "abso3bo", "Upper", "O3"), (25, "abso3ao", "Lower", "O3"),
                                (25, "abso3bo", "Upper", "O3"), (29, "absh2oo", "Lower", "H2O"),
                                (29, "absco2o", "Lower", "CO2")):
        add(b, arr, f"AbsorptionCoefficients{half_}Atmos", (1, 1, 1, ab[name]), (1, 1, 16, 1))
    return out


READS = _reads()


def read_blob(path=BLOB):
    """name -> Fortran-ordered array of the compiled-in table data."""
    raw = open(path, "rb").read()
    assert raw[:8] == b"RRTMGSWT"
    (nrec,) = struct.unpack_from("<I", raw, 12)
    data0 = 16 + nrec * 76
    out = {}
    for i in range(nrec):
        name, ndim, *rest = struct.unpack_from("<32si4i4iq", raw, 16 + 76 * i)
        ext, off = rest[4:8][:ndim], rest[8]
        n = int(np.prod(ext)) if ndim else 1
        a = np.frombuffer(raw, "<f8", n, data0 + 8 * off)
        out[name.rstrip(b"\0").decode()] = a.reshape(ext, order="F") if ndim else a.copy()
    return out


def build_arrays(seed=0):
    """File variable -> Fortran-ordered array holding the band data where read_nc reads it and
    junk elsewhere."""
    blob = read_blob()
    rng = np.random.default_rng(seed)
    arrays = {}
    for v, dims in VARS.items():
        shape = tuple(DIMS[d] for d in dims)
        arrays[v] = np.asfortranarray(-1.e3 * (1.0 + rng.random(shape)))
    for band, arr, var, start, count, rows in READS:
        src = blob[f"kg{band}_{arr}"]
        src = np.asarray(src).reshape(-1, order="F")
        if rows is not None:
            nrow = src.size // 16
            src = src.reshape(nrow, 16, order="F")[rows[0]:rows[1], :].reshape(-1, order="F")
        assert src.size == int(np.prod(count)), (band, arr, src.size, count)
        idx = tuple(slice(s - 1, s - 1 + c) for s, c in zip(start, count)) + (band - 16, 0)
        arrays[var][idx] = src.reshape(count, order="F")
    return arrays


def write(path, version=1, extras=False, seed=0):
    """Write the k-table file; `extras` adds attributes, variables of other types and a record
    variable the reader must step over."""
    from scipy.io import netcdf_file
    arrays = build_arrays(seed)
    f = netcdf_file(path, "w", version=version)
    try:
        if extras:
            f.title = "RRTMG_SW k-distribution test file"
            f.createDimension("time", None)
            f.createDimension("name_len", 7)
        for d, n in DIMS.items():
            f.createDimension(d, n)
        for v, dims in VARS.items():
            var = f.createVariable(v, "d", tuple(reversed(dims)))
            var[...] = arrays[v].T          # netCDF order is the Fortran order reversed
            if extras:
                var.units = "various"
                var.scale = np.array([1.0, 2.0])
        if extras:
            f.createVariable("AbsorberNames", "c", ("Absorber", "name_len"))[...] = \
                np.array([list(n.ljust(7)) for n in ABSORBERS], dtype="S1")
            f.createVariable("Temp", "f", ("T",))[...] = np.linspace(188.0, 317.6, 19)
            f.createVariable("BandNums", "i", ("band",))[...] = np.arange(16, 30)
            t = f.createVariable("time", "d", ("time",))
            t[:3] = [0.0, 1.0, 2.0]
            f.createVariable("flag", "h", ("time", "band"))[:3] = np.ones((3, 14))
    finally:
        f.close()
    return arrays


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("out")
    ap.add_argument("--version", type=int, default=1)
    ap.add_argument("--extras", action="store_true")
    a = ap.parse_args()
    write(a.out, a.version, a.extras)
    print(f"wrote {a.out}")


if __name__ == "__main__":
    main()
