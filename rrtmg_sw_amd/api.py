"""Python mirror of the reference's GCM-mode interface, running on the MI355X library.

Names, argument order and meaning follow the reference routines:
  rrtmg_sw_ini(cpdair)                          src/rrtmg_sw_init.f90:47
  rrtmg_sw(ncol, nlay, icld, iaer, play, ...)   src/rrtmg_sw_rad.f90:99-111
  rrtmg_sw_nomcica(ncol, nlay, icld, ...)       src/rrtmg_sw_rad.nomcica.f90:97-111 (the
                                                non-McICA driver's rrtmg_sw)
  mcica_subcol_sw(iplon, ncol, nlay, icld, ...) src/mcica_subcol_gen_sw.f90:181-185
  get_alpha(iplon, ncol, nlayers, icld, ...)    src/mcica_subcol_gen_sw.f90:66-67
plus mcica_rrtmg_sw(...), both calls fused on the GPU without the (112, ncol, nlay) arrays.

Arrays are either numpy arrays in the reference's Fortran layout (column-major, column index
fastest; the host-array path copies them to the GPU) or torch device tensors holding the same
memory layout, i.e. with the Fortran dimensions reversed (a (ncol, nlay) Fortran array is a
contiguous (nlay, ncol) tensor). Outputs come back in the same form as the inputs.
Errors raise RRTMGSWError carrying the reference's STOP message.
"""
import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import RRTMGSWError, check  # noqa: F401

NGPT, NBND, NAER = 112, 14, 6

try:  # torch is plumbing for device memory and streams only
    import torch
except Exception:  # pragma: no cover
    torch = None


def _is_dev(x):
    return torch is not None and isinstance(x, torch.Tensor) and x.is_cuda


def _ptr(x, n=None):
    dp = C.POINTER(C.c_double)
    if x is None:
        return None
    if _is_dev(x):
        if x.dtype != torch.float64 or not x.is_contiguous():
            raise ValueError("device arrays must be contiguous float64 tensors")
        if n is not None and x.numel() != n:
            raise ValueError(f"device array has {x.numel()} elements, expected {n}")
        return C.cast(C.c_void_p(x.data_ptr()), dp)
    a = np.asarray(x)
    if a.dtype != np.float64 or not (a.flags.f_contiguous or a.ndim <= 1):
        raise ValueError("host arrays must be Fortran-ordered float64")
    if n is not None and a.size != n:
        raise ValueError(f"host array has {a.size} elements, expected {n}")
    return a.ctypes.data_as(dp)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream) if torch is not None else None


def _new(like_dev, fshape):
    """Output array of Fortran shape `fshape` on the side of `like_dev`."""
    if like_dev:
        return torch.zeros(tuple(reversed(fshape)), dtype=torch.float64, device="cuda")
    return np.zeros(fshape, dtype=np.float64, order="F")


def to_device(a):
    """numpy Fortran-ordered array -> torch device tensor with the same memory layout."""
    a = np.asfortranarray(a, dtype=np.float64)
    return torch.from_numpy(np.ascontiguousarray(a.T)).cuda()


def to_host(t, fshape=None):
    """torch device tensor (reversed Fortran dims) -> numpy Fortran-ordered array."""
    a = t.detach().cpu().numpy()
    return np.asfortranarray(a.T) if fshape is None else np.asfortranarray(a.T.reshape(fshape))


def rrtmg_sw_ini(cpdair=1004.64, ktable_file=None):
    """Build and upload the 112-point tables (src/rrtmg_sw_init.f90:47). With `ktable_file`, the
    k-distribution and solar-source data come from that netCDF file, read as
    src/rrtmg_sw_read_nc.f90 reads rrtmg_sw.nc ("" names rrtmg_sw.nc in the working directory)."""
    if ktable_file is None:
        check(_lib.lib().rrtmgsw_init(C.c_double(cpdair)))
    else:
        check(_lib.lib().rrtmgsw_init_nc(C.c_double(cpdair), os.fsencode(ktable_file)))


def rrtmg_sw_finalize():
    """Free the device tables; the next rrtmg_sw_ini may take its tables from another source."""
    check(_lib.lib().rrtmgsw_finalize())


def host_tables(ktable_file=None):
    """The reduced tables built on the host (no GPU): the table image followed by four
    coefficients per g-point (rrtmgsw_host_tables)."""
    L = _lib.lib()
    path = None if ktable_file is None else os.fsencode(ktable_file)
    n = C.c_size_t(0)
    check(L.rrtmgsw_host_tables(path, None, 0, C.byref(n)))
    out = np.empty(n.value, dtype=np.float64)
    check(L.rrtmgsw_host_tables(path, out.ctypes.data_as(C.POINTER(C.c_double)), n.value,
                                C.byref(n)))
    return out


def nc_get_vara(path, name, start, count):
    """Hyperslab of a netCDF classic-format variable as float64, shape `count` (C order)."""
    start = (C.c_size_t * len(start))(*start)
    cnt = (C.c_size_t * len(count))(*count)
    out = np.empty(int(np.prod(count, dtype=np.int64)), dtype=np.float64)
    check(_lib.lib().rrtmgsw_nc_get_vara_double(os.fsencode(path), name.encode(), len(count),
                                                start, cnt,
                                                out.ctypes.data_as(C.POINTER(C.c_double))))
    return out.reshape(tuple(count))


def rrtmg_sw(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr,
             ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon,
             isolvar, inflgsw, iceflgsw, liqflgsw, cldfmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl,
             ciwpmcl, clwpmcl, reicmcl, relqmcl, tauaer, ssaaer, asmaer, ecaer,
             bndsolvar=None, indsolvar=None, solcycfrac=None, out=None):
    """rrtmg_sw (src/rrtmg_sw_rad.f90:99). Returns dict(swuflx, swdflx, swhr, swuflxc, swdflxc,
    swhrc, icld, iaer) -- icld / iaer as clamped by the reference (intent(inout)). `out` may
    hold the six output arrays of a previous call to write into."""
    dev = _is_dev(play)
    N, L, G = ncol, nlay, NGPT
    if out is None:
        out = {k: _new(dev, (N, L + 1)) for k in ("swuflx", "swdflx", "swuflxc", "swdflxc")}
        out["swhr"] = _new(dev, (N, L))
        out["swhrc"] = _new(dev, (N, L))
    ic, ia = C.c_int(icld), C.c_int(iaer)
    opt = [None if bndsolvar is None else np.ascontiguousarray(bndsolvar, dtype=np.float64),
           None if indsolvar is None else np.ascontiguousarray(indsolvar, dtype=np.float64),
           None if solcycfrac is None else np.array([solcycfrac], dtype=np.float64)]
    cl = icld >= 1 and icld <= 3
    args = [C.c_int(N), C.c_int(L), C.byref(ic), C.byref(ia),
            _ptr(play, N * L), _ptr(plev, N * (L + 1)), _ptr(tlay, N * L), _ptr(tlev), _ptr(tsfc),
            _ptr(h2ovmr, N * L), _ptr(o3vmr, N * L), _ptr(co2vmr, N * L), _ptr(ch4vmr, N * L),
            _ptr(n2ovmr), _ptr(o2vmr, N * L), _ptr(asdir, N), _ptr(asdif, N), _ptr(aldir, N),
            _ptr(aldif, N), _ptr(coszen, N), C.c_double(adjes), C.c_int(dyofyr), C.c_double(scon),
            C.c_int(isolvar), C.c_int(inflgsw), C.c_int(iceflgsw), C.c_int(liqflgsw),
            *[_ptr(x, G * N * L if cl else None) for x in (cldfmcl, taucmcl, ssacmcl, asmcmcl,
                                                          fsfcmcl, ciwpmcl, clwpmcl)],
            _ptr(reicmcl), _ptr(relqmcl), _ptr(tauaer), _ptr(ssaaer), _ptr(asmaer), _ptr(ecaer),
            _ptr(out["swuflx"]), _ptr(out["swdflx"]), _ptr(out["swhr"]), _ptr(out["swuflxc"]),
            _ptr(out["swdflxc"]), _ptr(out["swhrc"]), *[_ptr(o) for o in opt]]
    if dev:
        check(_lib.lib().rrtmgsw_sw_dev(*args, _stream()))
    else:
        check(_lib.lib().rrtmgsw_sw(*args))
    out["icld"], out["iaer"] = ic.value, ia.value
    return out


def rrtmg_sw_nomcica(ncol, nlay, icld, iaer, play, plev, tlay, tlev, tsfc, h2ovmr, o3vmr, co2vmr,
                     ch4vmr, n2ovmr, o2vmr, asdir, asdif, aldir, aldif, coszen, adjes, dyofyr, scon,
                     isolvar, inflgsw, iceflgsw, liqflgsw, cldfr, taucld, ssacld, asmcld, fsfcld,
                     cicewp, cliqwp, reice, reliq, tauaer, ssaaer, asmaer, ecaer,
                     bndsolvar=None, indsolvar=None, solcycfrac=None):
    """rrtmg_sw of the non-McICA driver (src/rrtmg_sw_rad.nomcica.f90:97): layer cloud fractions
    (clear or overcast: 'PARTIAL CLOUD NOT ALLOWED' otherwise) with cldprop_sw optics. Returns
    dict(swuflx, swdflx, swhr, swuflxc, swdflxc, swhrc, icld, iaer), icld / iaer as clamped."""
    dev = _is_dev(play)
    N, L = ncol, nlay
    out = {k: _new(dev, (N, L + 1)) for k in ("swuflx", "swdflx", "swuflxc", "swdflxc")}
    out["swhr"] = _new(dev, (N, L))
    out["swhrc"] = _new(dev, (N, L))
    ic, ia = C.c_int(icld), C.c_int(iaer)
    opt = [None if bndsolvar is None else np.ascontiguousarray(bndsolvar, dtype=np.float64),
           None if indsolvar is None else np.ascontiguousarray(indsolvar, dtype=np.float64),
           None if solcycfrac is None else np.array([solcycfrac], dtype=np.float64)]
    cl = icld != 0          # icld outside 0..3 is clamped to 2: clouds are read unless icld = 0
    nl = N * L if cl else None
    nb = NBND * N * L if cl else None
    args = [C.c_int(N), C.c_int(L), C.byref(ic), C.byref(ia),
            _ptr(play, N * L), _ptr(plev, N * (L + 1)), _ptr(tlay, N * L), _ptr(tlev), _ptr(tsfc),
            _ptr(h2ovmr, N * L), _ptr(o3vmr, N * L), _ptr(co2vmr, N * L), _ptr(ch4vmr, N * L),
            _ptr(n2ovmr), _ptr(o2vmr, N * L), _ptr(asdir, N), _ptr(asdif, N), _ptr(aldir, N),
            _ptr(aldif, N), _ptr(coszen, N), C.c_double(adjes), C.c_int(dyofyr), C.c_double(scon),
            C.c_int(isolvar), C.c_int(inflgsw), C.c_int(iceflgsw), C.c_int(liqflgsw),
            _ptr(cldfr, nl), *[_ptr(x, nb) for x in (taucld, ssacld, asmcld, fsfcld)],
            *[_ptr(x, nl) for x in (cicewp, cliqwp, reice, reliq)],
            _ptr(tauaer), _ptr(ssaaer), _ptr(asmaer), _ptr(ecaer),
            _ptr(out["swuflx"]), _ptr(out["swdflx"]), _ptr(out["swhr"]), _ptr(out["swuflxc"]),
            _ptr(out["swdflxc"]), _ptr(out["swhrc"]), *[_ptr(o) for o in opt]]
    if dev:
        check(_lib.lib().rrtmgsw_sw_nomcica_dev(*args, _stream()))
    else:
        check(_lib.lib().rrtmgsw_sw_nomcica(*args))
    out["icld"], out["iaer"] = ic.value, ia.value
    return out


def mcica_subcol_sw(iplon, ncol, nlay, icld, permuteseed, irng, play, cldfrac, ciwp, clwp, rei,
                    rel, tauc, ssac, asmc, fsfc, alpha, out=None):
    """mcica_subcol_sw (src/mcica_subcol_gen_sw.f90:181). Returns dict of the nine outputs
    (cldfmcl, ciwpmcl, clwpmcl, reicmcl, relqmcl, taucmcl, ssacmcl, asmcmcl, fsfcmcl) and
    irng as written back. `out` may hold the nine arrays of a previous call to write into."""
    dev = _is_dev(play)
    N, L, G = ncol, nlay, NGPT
    if out is None:
        out = {k: _new(dev, (G, N, L)) for k in ("cldfmcl", "ciwpmcl", "clwpmcl", "taucmcl",
                                                  "ssacmcl", "asmcmcl", "fsfcmcl")}
        out["reicmcl"] = _new(dev, (N, L))
        out["relqmcl"] = _new(dev, (N, L))
    ir = C.c_int(irng)
    args = [C.c_int(iplon), C.c_int(N), C.c_int(L), C.c_int(icld), C.c_int(permuteseed),
            C.byref(ir), _ptr(play, N * L), _ptr(cldfrac, N * L), _ptr(ciwp, N * L),
            _ptr(clwp, N * L), _ptr(rei, N * L), _ptr(rel, N * L), _ptr(tauc, NBND * N * L),
            _ptr(ssac, NBND * N * L), _ptr(asmc, NBND * N * L), _ptr(fsfc, NBND * N * L),
            _ptr(alpha, N * L)] + [_ptr(out[k]) for k in _lib.MC_OUT]
    if dev:
        check(_lib.lib().rrtmgsw_mcica_subcol_dev(*args, _stream()))
    else:
        check(_lib.lib().rrtmgsw_mcica_subcol(*args))
    out["irng"] = ir.value
    return out


def get_alpha(iplon, ncol, nlayers, icld, idcor, decorr_con, dz, lat, juldat, cldfrac):
    """get_alpha (src/mcica_subcol_gen_sw.f90:66); host arrays."""
    alpha = np.zeros((ncol, nlayers), order="F")
    check(_lib.lib().rrtmgsw_get_alpha(iplon, ncol, nlayers, icld, idcor, C.c_double(decorr_con),
                                       _ptr(dz), _ptr(np.ascontiguousarray(lat, dtype=np.float64)),
                                       juldat, _ptr(cldfrac), _ptr(alpha)))
    return alpha


# setcoef_sw's outputs in the reference's dummy order (src/rrtmg_sw_setcoef.f90:49-53), as
# rrtmgsw_setcoef_dev writes them, then laysolfr; True marks the int32 ones
SETCOEF_OUT = [("laytrop", True), ("laylow", True), ("jp", True), ("jt", True), ("jt1", True),
               ("colch4", False), ("colco2", False), ("colh2o", False), ("colmol", False),
               ("colo2", False), ("colo3", False), ("fac00", False), ("fac01", False),
               ("fac10", False), ("fac11", False), ("selffac", False), ("selffrac", False),
               ("indself", True), ("forfac", False), ("forfrac", False), ("indfor", True),
               ("laysolfr", True)]


def setcoef_sw(ncol, nlay, play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr):
    """setcoef_sw (src/rrtmg_sw_setcoef.f90:49-305) for every column on the GPU, fed from
    inatm_sw's layer amounts of the given device profiles (torch tensors in reversed Fortran
    dims). Returns device tensors: laytrop, laylow (ncol,); laysolfr (14, ncol) (Fortran
    (ncol, 14)); the rest (nlay, ncol) (Fortran (ncol, nlay)); indices 1-based int32."""
    ip = C.POINTER(C.c_int)
    out = {}
    for name, is_int in SETCOEF_OUT:
        shape = (ncol,) if name in ("laytrop", "laylow") else ((NBND, ncol) if name == "laysolfr"
                                                               else (nlay, ncol))
        out[name] = torch.zeros(shape, dtype=torch.int32 if is_int else torch.float64, device="cuda")
    args = [C.c_int(ncol), C.c_int(nlay)]
    args += [_ptr(x) for x in (play, plev, tlay, h2ovmr, o3vmr, co2vmr, ch4vmr, o2vmr)]
    for name, is_int in SETCOEF_OUT:
        t = out[name]
        args.append(C.cast(C.c_void_p(t.data_ptr()), ip) if is_int else _ptr(t))
    args.append(_stream())
    check(_lib.lib().rrtmgsw_setcoef_dev(*args))
    return out


PRECISION = {"fp64": 0, "mixed": 1, "exact": 2}


def _solvar(bndsolvar, indsolvar, solcycfrac):
    """The optional solar-variability arguments as host arrays (read on the host by every entry
    point), or None."""
    return [None if bndsolvar is None else np.ascontiguousarray(bndsolvar, dtype=np.float64),
            None if indsolvar is None else np.ascontiguousarray(indsolvar, dtype=np.float64),
            None if solcycfrac is None else np.array([solcycfrac], dtype=np.float64)]


def mcica_rrtmg_sw(cols, dev_arrays, precision="fp64", out=None, bndsolvar=None, indsolvar=None,
                   solcycfrac=None, noalloc=False):
    """mcica_subcol_sw + rrtmg_sw fused on the GPU for the workload `cols` (a
    workloads.Columns) whose arrays are already resident: `dev_arrays` maps input names to
    device tensors (see device_inputs). bndsolvar (14), indsolvar (2) and solcycfrac are the
    reference's optional solar-variability arguments, on the host. noalloc selects the entry
    point that allocates no device memory (rrtmgsw_mcica_sw_dev_noalloc). Returns the output
    tensors."""
    N, L = cols.ncol, cols.nlay
    d = dev_arrays
    if out is None:
        out = {k: _new(True, (N, L + 1)) for k in ("swuflx", "swdflx", "swuflxc", "swdflxc")}
        out["swhr"] = _new(True, (N, L))
        out["swhrc"] = _new(True, (N, L))
    ir = C.c_int(cols.irng)
    args = [C.c_int(N), C.c_int(L), C.c_int(cols.icld), C.c_int(cols.iaer),
            C.c_int(cols.permuteseed), C.byref(ir)]
    args += [_ptr(d.get(k)) for k in _lib.SW_PROFILE]
    args += [C.c_double(cols.adjes), C.c_int(cols.dyofyr), C.c_double(cols.scon),
             C.c_int(cols.isolvar), C.c_int(cols.inflgsw), C.c_int(cols.iceflgsw),
             C.c_int(cols.liqflgsw)]
    args += [_ptr(d.get(k)) for k in _lib.FUSED_CLD]
    args += [_ptr(d.get(k)) for k in _lib.SW_AER]
    args += [_ptr(out[k]) for k in _lib.SW_OUT]
    sv = _solvar(bndsolvar, indsolvar, solcycfrac)
    args += [_ptr(x) for x in sv] + [C.c_int(PRECISION[precision]), _stream()]
    L = _lib.lib()
    check((L.rrtmgsw_mcica_sw_dev_noalloc if noalloc else L.rrtmgsw_mcica_sw_dev)(*args))
    return out


def _fused_inputs(cols, d, irng):
    args = [C.c_int(cols.ncol), C.c_int(cols.nlay), C.c_int(cols.icld), C.c_int(cols.iaer),
            C.c_int(cols.permuteseed), C.byref(irng)]
    args += [_ptr(d.get(k)) for k in _lib.SW_PROFILE]
    args += [C.c_double(cols.adjes), C.c_int(cols.dyofyr), C.c_double(cols.scon),
             C.c_int(cols.isolvar), C.c_int(cols.inflgsw), C.c_int(cols.iceflgsw),
             C.c_int(cols.liqflgsw)]
    args += [_ptr(d.get(k)) for k in _lib.FUSED_CLD]
    args += [_ptr(d.get(k)) for k in _lib.SW_AER]
    return args


ADJ_INPUTS = ("tlay", "h2ovmr", "o3vmr", "clwp", "ciwp")


def mcica_rrtmg_sw_adjoint(cols, dev_arrays, seed_toa, seed_hr, out=None, ckpt=0, bndsolvar=None,
                           indsolvar=None, solcycfrac=None):
    """Adjoint of the fused call (exact-exponential FP64 forward, masks fixed):
    J = sum seed_toa * swuflx(:, top) + sum seed_hr * swhr. seed_toa is a device tensor of
    shape (ncol,), seed_hr one of Fortran shape (ncol, nlay) (i.e. torch shape (nlay, ncol)).
    Returns device tensors dJ/d{tlay, h2ovmr, o3vmr, clwp, ciwp} (Fortran (ncol, nlay)) and
    dJ/d albedo ("alb", Fortran (ncol, 4): asdir, asdif, aldir, aldif). ckpt = 0 stores every
    layer's adding-method state, ckpt = K > 0 one per K layers (recomputing segments)."""
    N, L = cols.ncol, cols.nlay
    if out is None:
        out = {k: _new(True, (N, L)) for k in ADJ_INPUTS}
        out["alb"] = _new(True, (N, 4))
    ir = C.c_int(cols.irng)
    args = _fused_inputs(cols, dev_arrays, ir)
    sv = _solvar(bndsolvar, indsolvar, solcycfrac)
    args += [_ptr(x) for x in sv] + [_ptr(seed_toa, N), _ptr(seed_hr, N * L)]
    args += [_ptr(out[k]) for k in ADJ_INPUTS] + [_ptr(out["alb"]), C.c_int(ckpt), _stream()]
    check(_lib.lib().rrtmgsw_mcica_sw_adjoint_dev(*args))
    return out


def device_inputs(cols):
    """Copy every array input of `cols` to the GPU (reversed-dims tensors)."""
    names = _lib.SW_PROFILE + _lib.FUSED_CLD + _lib.SW_AER
    return {k: to_device(getattr(cols, k)) for k in names if getattr(cols, k) is not None}
