"""ctypes bindings for the checkers: the C oracle (oracle/liborc.so), its operation-counting build
(oracle/liborc_count.so) and the compiled reference (oracle/_ref/librrtmg_ref.so). Test
infrastructure only."""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BLOB = os.path.join(ROOT, "rrtmg_sw_amd", "data", "rrtmg_sw_tables.bin")
ORC_PATH = os.path.join(ROOT, "oracle", "liborc.so")
REF_PATH = os.path.join(ROOT, "oracle", "_ref", "librrtmg_ref.so")
REF_NM_PATH = os.path.join(ROOT, "oracle", "_ref", "librrtmg_ref_nomcica.so")
CPDAIR = 1004.64
NGPT, NBND = 112, 14

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.ndim <= 1), a.flags
    return a.ctypes.data_as(_dp)


def _ip_of(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


COUNT_PATH = os.path.join(ROOT, "oracle", "liborc_count.so")

_orc = None
_cnt = None
_ref = None
_ref_nm = None


def orc():
    global _orc
    if _orc is None:
        _orc = _load_orc(ORC_PATH)
    return _orc


def orc_count():
    """The oracle's operation-counting build (oracle/orc_count.cpp): the same interface, plus
    orc_fp_reset / orc_fp_counts."""
    global _cnt
    if _cnt is None:
        _cnt = _load_orc(COUNT_PATH)
        _cnt.orc_fp_counts.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    return _cnt


def fp_counts():
    """(+-*/ and sqrt, exp/log/pow) counted by the counting build since its last reset."""
    a, b = C.c_uint64(0), C.c_uint64(0)
    orc_count().orc_fp_counts(C.byref(a), C.byref(b))
    return a.value, b.value


def _lib(which):
    return orc_count() if which == "count" else orc()


def _load_orc(path):
    lib = C.CDLL(path)
    lib.orc_init.argtypes = [C.c_char_p, C.c_double]
    lib.orc_last_error.restype = C.c_char_p
    lib.orc_exp_tbl.restype = C.c_double
    lib.orc_exp_tbl.argtypes = [C.c_int]
    lib.orc_table.argtypes = [C.c_char_p, _dp]
    lib.orc_kiss_step.restype = C.c_double
    lib.orc_mt_first_int.restype = C.c_uint
    lib.orc_mt_first_int.argtypes = [C.c_int, C.c_long]
    lib.orc_mt_reals.argtypes = [C.c_int, C.c_long, _dp]
    rc = lib.orc_init(BLOB.encode(), CPDAIR)
    if rc:
        raise RuntimeError(lib.orc_last_error().decode())
    return lib


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_PATH)
        lib.ref_init.argtypes = [C.c_double]
        lib.ref_table.restype = C.c_int
        lib.ref_table.argtypes = [C.c_int, C.c_int, _dp]
        lib.ref_init(CPDAIR)
        _ref = lib
    return _ref


def ref_nomcica_available():
    return os.path.exists(REF_NM_PATH)


def ref_nomcica():
    """The compiled reference's non-McICA driver (its own library: the module names clash)."""
    global _ref_nm
    if _ref_nm is None:
        lib = C.CDLL(REF_NM_PATH)
        lib.ref_nm_init.argtypes = [C.c_double]
        lib.ref_nm_init(CPDAIR)
        _ref_nm = lib
    return _ref_nm


def run_nomcica(cols, which="orc", icld=None, iaer=None, solvar=None):
    """rrtmg_sw of the non-McICA driver on `cols` (layer clouds cols.cldfrac, cols.tauc, ...);
    solvar = (bndsolvar, indsolvar, solcycfrac) or None. Returns outputs and written-back flags;
    raises RuntimeError with the STOP message for the oracle's non-zero returns."""
    N, L = cols.ncol, cols.nlay
    out = {k: _out((N, L + 1)) for k in ("swuflx", "swdflx", "swuflxc", "swdflxc")}
    out["swhr"] = _out((N, L))
    out["swhrc"] = _out((N, L))
    ic = C.c_int(cols.icld if icld is None else icld)
    ia = C.c_int(cols.iaer if iaer is None else iaer)
    common = [C.c_int(N), C.c_int(L), C.byref(ic), C.byref(ia), _p(cols.play), _p(cols.plev),
              _p(cols.tlay), _p(cols.tlev), _p(cols.tsfc), _p(cols.h2ovmr), _p(cols.o3vmr),
              _p(cols.co2vmr), _p(cols.ch4vmr), _p(cols.n2ovmr), _p(cols.o2vmr), _p(cols.asdir),
              _p(cols.asdif), _p(cols.aldir), _p(cols.aldif), _p(cols.coszen),
              C.c_double(cols.adjes), C.c_int(cols.dyofyr), C.c_double(cols.scon),
              C.c_int(cols.isolvar), C.c_int(cols.inflgsw), C.c_int(cols.iceflgsw),
              C.c_int(cols.liqflgsw), _p(cols.cldfrac), _p(cols.tauc), _p(cols.ssac),
              _p(cols.asmc), _p(cols.fsfc), _p(cols.ciwp), _p(cols.clwp), _p(cols.rei),
              _p(cols.rel), _p(cols.tauaer), _p(cols.ssaaer), _p(cols.asmaer), _p(cols.ecaer),
              _p(out["swuflx"]), _p(out["swdflx"]), _p(out["swhr"]), _p(out["swuflxc"]),
              _p(out["swdflxc"]), _p(out["swhrc"])]
    sv = None
    if solvar is not None:
        sv = [np.ascontiguousarray(solvar[0], dtype=np.float64),
              np.ascontiguousarray(solvar[1], dtype=np.float64),
              np.array([solvar[2]], dtype=np.float64)]
    if which in ("orc", "count"):
        lib = _lib(which)
        rc = lib.orc_rrtmg_sw_nomcica(*common, *([None] * 3 if sv is None else [_p(x) for x in sv]))
        if rc:
            raise RuntimeError(lib.orc_last_error().decode())
    else:
        if sv is None:
            sv = [np.ones(NBND), np.ones(2), np.zeros(1)]
            has = 0
        else:
            has = 1
        ref_nomcica().ref_rrtmg_sw_nomcica(*common, C.c_int(has), *[_p(x) for x in sv])
    return out, ic.value, ia.value


def _out(shape):
    return np.zeros(shape, dtype=np.float64, order="F")


def mcica_outputs(cols):
    G, N, L = NGPT, cols.ncol, cols.nlay
    return {k: _out((G, N, L)) for k in ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl", "ssacmcl",
                                          "asmcmcl", "fsfcmcl")} | {
        "reicmcl": _out((N, L)), "relqmcl": _out((N, L))}


def run_mcica(cols, which="orc"):
    """mcica_subcol_sw on `cols`; returns the output dict and the written-back irng."""
    o = mcica_outputs(cols)
    irng = C.c_int(cols.irng)
    args = [C.c_int(1), C.c_int(cols.ncol), C.c_int(cols.nlay), C.c_int(cols.icld),
            C.c_int(cols.permuteseed), C.byref(irng), _p(cols.play), _p(cols.cldfrac),
            _p(cols.ciwp), _p(cols.clwp), _p(cols.rei), _p(cols.rel), _p(cols.tauc),
            _p(cols.ssac), _p(cols.asmc), _p(cols.fsfc), _p(cols.alpha),
            _p(o["cldfmcl"]), _p(o["ciwpmcl"]), _p(o["clwpmcl"]), _p(o["reicmcl"]),
            _p(o["relqmcl"]), _p(o["taucmcl"]), _p(o["ssacmcl"]), _p(o["asmcmcl"]),
            _p(o["fsfcmcl"])]
    if which in ("orc", "count"):
        rc = _lib(which).orc_mcica_subcol_sw(*args)
        if rc:
            raise RuntimeError(_lib(which).orc_last_error().decode())
    else:
        ref().ref_mcica_subcol_sw(*args)
    return o, irng.value


SETCOEF_INT = ["jp", "jt", "jt1", "indself", "indfor"]
SETCOEF_DBL = ["colh2o", "colco2", "colo3", "colch4", "colo2", "colmol", "fac00", "fac01",
               "fac10", "fac11", "selffac", "selffrac", "forfac", "forfrac"]


def run_setcoef(ncol, nlay, play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr):
    """The oracle's setcoef_sw for every column, fed from inatm_sw's layer amounts (Fortran
    (ncol, nlay) / (ncol, nlay+1) float64 arrays). Returns laytrop, laylow (ncol), laysolfr
    (ncol, 14) and the per-layer arrays (ncol, nlay), in the reference's 1-based indices."""
    N, L = ncol, nlay
    iout = np.zeros((2 + NBND + 5 * L) * N, dtype=np.int32)
    dout = np.zeros(14 * N * L, dtype=np.float64)
    arrs = [np.asfortranarray(a, dtype=np.float64) for a in (play, plev, tlay, h2ovmr, o3vmr,
                                                             co2vmr, ch4vmr, o2vmr)]
    orc().orc_setcoef_gcm(C.c_int(N), C.c_int(L), *[_p(a) for a in arrs], _ip_of(iout),
                          _p(dout))
    o = {"laytrop": iout[:N].copy(), "laylow": iout[N:2 * N].copy(),
         "laysolfr": iout[2 * N:(2 + NBND) * N].reshape((N, NBND), order="F").copy()}
    base = (2 + NBND) * N
    for k, name in enumerate(SETCOEF_INT):
        o[name] = iout[base + k * N * L: base + (k + 1) * N * L].reshape((N, L), order="F").copy()
    for k, name in enumerate(SETCOEF_DBL):
        o[name] = dout[k * N * L:(k + 1) * N * L].reshape((N, L), order="F").copy()
    return o


def empty_mcica(cols):
    G, N, L = NGPT, cols.ncol, cols.nlay
    o = {k: _out((G, N, L)) for k in ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl", "asmcmcl",
                                       "fsfcmcl")}
    o["ssacmcl"] = np.ones((G, N, L), order="F")
    o["reicmcl"] = _out((N, L))
    o["relqmcl"] = _out((N, L))
    return o


def run_rrtmg(cols, mc=None, which="orc", bndsolvar=None, indsolvar=None, solcycfrac=None,
              icld=None, iaer=None, bands=False, idelm=1):
    """rrtmg_sw on `cols` with McICA arrays `mc`; returns outputs and written-back flags."""
    N, L = cols.ncol, cols.nlay
    if mc is None:
        mc = empty_mcica(cols)
    out = {k: _out((N, L + 1)) for k in ("swuflx", "swdflx", "swuflxc", "swdflxc")}
    out["swhr"] = _out((N, L))
    out["swhrc"] = _out((N, L))
    ic = C.c_int(cols.icld if icld is None else icld)
    ia = C.c_int(cols.iaer if iaer is None else iaer)
    bnd = None if bndsolvar is None else np.ascontiguousarray(bndsolvar, dtype=np.float64)
    ind = None if indsolvar is None else np.ascontiguousarray(indsolvar, dtype=np.float64)
    frac = None if solcycfrac is None else np.array([solcycfrac], dtype=np.float64)
    common = [C.c_int(N), C.c_int(L), C.byref(ic), C.byref(ia), _p(cols.play), _p(cols.plev),
              _p(cols.tlay), _p(cols.tlev), _p(cols.tsfc), _p(cols.h2ovmr), _p(cols.o3vmr),
              _p(cols.co2vmr), _p(cols.ch4vmr), _p(cols.n2ovmr), _p(cols.o2vmr), _p(cols.asdir),
              _p(cols.asdif), _p(cols.aldir), _p(cols.aldif), _p(cols.coszen),
              C.c_double(cols.adjes), C.c_int(cols.dyofyr), C.c_double(cols.scon),
              C.c_int(cols.isolvar), C.c_int(cols.inflgsw), C.c_int(cols.iceflgsw),
              C.c_int(cols.liqflgsw), _p(mc["cldfmcl"]), _p(mc["taucmcl"]), _p(mc["ssacmcl"]),
              _p(mc["asmcmcl"]), _p(mc["fsfcmcl"]), _p(mc["ciwpmcl"]), _p(mc["clwpmcl"]),
              _p(mc["reicmcl"]), _p(mc["relqmcl"]), _p(cols.tauaer), _p(cols.ssaaer),
              _p(cols.asmaer), _p(cols.ecaer), _p(out["swuflx"]), _p(out["swdflx"]),
              _p(out["swhr"]), _p(out["swuflxc"]), _p(out["swdflxc"]), _p(out["swhrc"])]
    if which in ("orc", "count"):
        bu = bd = dird = None
        if bands:
            bu, bd = _out((N, L + 1, NBND)), _out((N, L + 1, NBND))
            dird = _out((N, L + 1))
        lib = _lib(which)
        rc = lib.orc_rrtmg_sw_ext(*common, _p(bnd), _p(ind), _p(frac), _p(bu), _p(bd),
                                  _p(dird), C.c_int(idelm))
        if rc:
            raise RuntimeError(lib.orc_last_error().decode())
        if bands:
            out["bandu"], out["bandd"], out["dirdflx"] = bu, bd, dird
    else:
        dummy_b = np.zeros(NBND)
        dummy_i = np.zeros(2)
        dummy_f = np.zeros(1)
        ref().ref_rrtmg_sw(*common,
                           C.c_int(bnd is not None), _p(bnd if bnd is not None else dummy_b),
                           C.c_int(ind is not None), _p(ind if ind is not None else dummy_i),
                           C.c_int(frac is not None), _p(frac if frac is not None else dummy_f))
    out["icld"], out["iaer"] = ic.value, ia.value
    return out


TABLE_KINDS = ["ka", "kb", "selfref", "forref", "sfluxref", "irradnce", "facbrght", "snsptdrk",
               "rayl", "rayla", "raylb", "abso3a", "abso3b", "absch4", "absh2o", "absco2"]


def orc_table(name):
    n = orc().orc_table(name.encode(), None)
    if n <= 0:
        return None
    a = np.zeros(n)
    orc().orc_table(name.encode(), _p(a))
    return a


def ref_table(band, kind):
    buf = np.zeros(20000)
    n = ref().ref_table(band, TABLE_KINDS.index(kind) + 1, _p(buf))
    return buf[:n].copy() if n > 0 else None
