"""ctypes binding of librrtmgsw_hip.so (the C-ABI declared in include/rrtmgsw.h).

The library is built in-tree by rrtmg_sw_amd.build; importing this module without it raises
immediately -- there is no CPU fallback.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librrtmgsw_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p

# Argument lists, in the order of the reference's Fortran dummies.
SW_PROFILE = ["play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr",
              "n2ovmr", "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen"]
SW_MCICA = ["cldfmcl", "taucmcl", "ssacmcl", "asmcmcl", "fsfcmcl", "ciwpmcl", "clwpmcl",
            "reicmcl", "relqmcl"]
SW_AER = ["tauaer", "ssaaer", "asmaer", "ecaer"]
SW_OUT = ["swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"]
MC_IN = ["play", "cldfrac", "ciwp", "clwp", "rei", "rel", "tauc", "ssac", "asmc", "fsfc", "alpha"]
MC_OUT = ["cldfmcl", "ciwpmcl", "clwpmcl", "reicmcl", "relqmcl", "taucmcl", "ssacmcl",
          "asmcmcl", "fsfcmcl"]
FUSED_CLD = ["cldfrac", "ciwp", "clwp", "rei", "rel", "tauc", "ssac", "asmc", "fsfc", "alpha"]

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m rrtmg_sw_amd.build`")
    L = C.CDLL(LIB_PATH)
    L.rrtmgsw_init.argtypes = [C.c_double]
    L.rrtmgsw_init_nc.argtypes = [C.c_double, C.c_char_p]
    L.rrtmgsw_host_tables.argtypes = [C.c_char_p, _dp, C.c_size_t, C.POINTER(C.c_size_t)]
    L.rrtmgsw_nc_get_vara_double.argtypes = [C.c_char_p, C.c_char_p, C.c_int,
                                             C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), _dp]
    L.rrtmgsw_finalize.argtypes = []
    L.rrtmgsw_last_error.restype = C.c_char_p
    sw = ([C.c_int, C.c_int, _ip, _ip] + [_dp] * 16 + [C.c_double, C.c_int, C.c_double, C.c_int,
          C.c_int, C.c_int, C.c_int] + [_dp] * 9 + [_dp] * 4 + [_dp] * 6 + [_dp] * 3)
    L.rrtmgsw_sw.argtypes = sw
    L.rrtmgsw_sw_dev.argtypes = sw + [_vp]
    L.rrtmgsw_sw_nomcica.argtypes = sw          # same dummy kinds, layer clouds in place of McICA
    L.rrtmgsw_sw_nomcica_dev.argtypes = sw + [_vp]
    mc = [C.c_int] * 5 + [_ip] + [_dp] * 11 + [_dp] * 9
    L.rrtmgsw_mcica_subcol.argtypes = mc
    L.rrtmgsw_mcica_subcol_dev.argtypes = mc + [_vp]
    L.rrtmgsw_get_alpha.argtypes = ([C.c_int] * 5 + [C.c_double, _dp, _dp, C.c_int, _dp, _dp])
    L.rrtmgsw_mcica_sw_dev.argtypes = ([C.c_int] * 5 + [_ip] + [_dp] * 16 +
                                       [C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                        C.c_int] + [_dp] * 10 + [_dp] * 4 + [_dp] * 6 + [_dp] * 3 +
                                       [C.c_int, _vp])
    L.rrtmgsw_mcica_sw_dev_noalloc.argtypes = L.rrtmgsw_mcica_sw_dev.argtypes
    L.rrtmgsw_mcica_sw_adjoint_dev.argtypes = (
        [C.c_int] * 5 + [_ip] + [_dp] * 16 +
        [C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int] + [_dp] * 10 +
        [_dp] * 4 + [_dp] * 3 + [_dp] * 2 + [_dp] * 6 + [C.c_int, _vp])
    L.rrtmgsw_setcoef_dev.argtypes = ([C.c_int, C.c_int] + [_dp] * 8 + [_ip] * 5 + [_dp] * 12 +
                                      [_ip, _dp, _dp, _ip, _ip, _vp])
    L.rrtmgsw_adjoint_work_bytes.argtypes = [C.c_int, C.c_int]
    L.rrtmgsw_adjoint_work_bytes.restype = C.c_size_t
    _lib = L
    return L


EXPORTED = ["rrtmgsw_init", "rrtmgsw_init_nc", "rrtmgsw_host_tables", "rrtmgsw_nc_get_vara_double",
            "rrtmgsw_finalize", "rrtmgsw_last_error", "rrtmgsw_sw",
            "rrtmgsw_sw_dev", "rrtmgsw_sw_nomcica", "rrtmgsw_sw_nomcica_dev", "rrtmgsw_column_sw",
            "rrtmgsw_column_mcica_sw",
            "rrtmgsw_mcica_subcol", "rrtmgsw_mcica_subcol_dev",
            "rrtmgsw_get_alpha", "rrtmgsw_setcoef_dev", "rrtmgsw_mcica_sw_dev",
            "rrtmgsw_mcica_sw_dev_noalloc",
            "rrtmgsw_mcica_sw_adjoint_dev", "rrtmgsw_adjoint_work_bytes"]


def check(rc):
    if rc != 0:
        msg = lib().rrtmgsw_last_error().decode()
        raise RRTMGSWError(rc, msg)


class RRTMGSWError(RuntimeError):
    """A non-zero return of the C-ABI: the reference would STOP with `message`."""

    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message
