"""HIP library's non-McICA driver (rrtmgsw_sw_nomcica / _dev: layer cloud fractions, cldprop_sw
optics on every g-point of the band) against the C oracle on the same seeded inputs. Bars as for
the McICA path: fluxes within 0.1 W/m2, heating rates within 0.01 K/day, the median flux
difference at the rounding level; written-back flags and error codes equal."""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload
from test_gpu_parity import FLUX, HR, assert_close
from test_oracle_nomcica import CASES

pytestmark = pytest.mark.gpu


def nm_args(cols, icld=None, iaer=None):
    return dict(ncol=cols.ncol, nlay=cols.nlay, icld=cols.icld if icld is None else icld,
                iaer=cols.iaer if iaer is None else iaer, play=cols.play, plev=cols.plev,
                tlay=cols.tlay, tlev=cols.tlev, tsfc=cols.tsfc, h2ovmr=cols.h2ovmr,
                o3vmr=cols.o3vmr, co2vmr=cols.co2vmr, ch4vmr=cols.ch4vmr, n2ovmr=cols.n2ovmr,
                o2vmr=cols.o2vmr, asdir=cols.asdir, asdif=cols.asdif, aldir=cols.aldir,
                aldif=cols.aldif, coszen=cols.coszen, adjes=cols.adjes, dyofyr=cols.dyofyr,
                scon=cols.scon, isolvar=cols.isolvar, inflgsw=cols.inflgsw,
                iceflgsw=cols.iceflgsw, liqflgsw=cols.liqflgsw, cldfr=cols.cldfrac,
                taucld=cols.tauc, ssacld=cols.ssac, asmcld=cols.asmc, fsfcld=cols.fsfc,
                cicewp=cols.ciwp, cliqwp=cols.clwp, reice=cols.rei, reliq=cols.rel,
                tauaer=cols.tauaer, ssaaer=cols.ssaaer, asmaer=cols.asmaer, ecaer=cols.ecaer)


@pytest.mark.parametrize("name", sorted(CASES))
def test_nomcica_host_abi(gpu_lib, name):
    cols = CASES[name]
    got = gpu_lib.rrtmg_sw_nomcica(**nm_args(cols))
    ref, ic, ia = ob.run_nomcica(cols)
    assert (got["icld"], got["iaer"]) == (ic, ia)
    print("nomcica", name, {k: float(np.abs(got[k] - ref[k]).max()) for k in FLUX + HR}, flush=True)
    assert_close(got, ref, name)


@pytest.mark.parametrize("icld,iaer", [(0, 0), (7, 3)])
def test_nomcica_flag_clamps(gpu_lib, icld, iaer):
    cols = workload("overcast8192x72", ncol=8)
    got = gpu_lib.rrtmg_sw_nomcica(**nm_args(cols, icld=icld, iaer=iaer))
    ref, ic, ia = ob.run_nomcica(cols, icld=icld, iaer=iaer)
    assert (got["icld"], got["iaer"]) == (ic, ia) == ((0, 0) if icld == 0 else (2, 0))
    assert_close(got, ref, f"icld={icld}")


def test_nomcica_partial_cloud_error(gpu_lib):
    cols = workload("overcast8192x72", ncol=8)
    f = cols.cldfrac.copy()
    f[5, np.argmax(f[5] > 0)] = 0.5
    cols.cldfrac = np.asfortranarray(f)
    with pytest.raises(gpu_lib.RRTMGSWError, match="PARTIAL CLOUD NOT ALLOWED") as e:
        gpu_lib.rrtmg_sw_nomcica(**nm_args(cols))
    assert e.value.code == 6


def test_nomcica_overcast8192x72_device(gpu_lib):
    """The full 8192 x 72 workload on device arrays; the first and last 256 columns against the
    oracle (columns are independent, so a subset of the full call is exact to compare)."""
    import torch
    from rrtmg_sw_amd import api
    cols = workload("overcast8192x72")
    d = api.device_inputs(cols)
    out = api.rrtmg_sw_nomcica(**{**nm_args(cols), **{k: d[n] for k, n in (
        ("play", "play"), ("plev", "plev"), ("tlay", "tlay"), ("h2ovmr", "h2ovmr"),
        ("o3vmr", "o3vmr"), ("co2vmr", "co2vmr"), ("ch4vmr", "ch4vmr"), ("o2vmr", "o2vmr"),
        ("asdir", "asdir"), ("asdif", "asdif"), ("aldir", "aldir"), ("aldif", "aldif"),
        ("coszen", "coszen"), ("cldfr", "cldfrac"), ("taucld", "tauc"), ("ssacld", "ssac"),
        ("asmcld", "asmc"), ("fsfcld", "fsfc"), ("cicewp", "ciwp"), ("cliqwp", "clwp"),
        ("reice", "rei"), ("reliq", "rel"))}, "tlev": None, "tsfc": None, "n2ovmr": None,
        "tauaer": None, "ssaaer": None, "asmaer": None, "ecaer": None})
    torch.cuda.synchronize()
    got = {k: api.to_host(v) for k, v in out.items() if k not in ("icld", "iaer")}
    for sl in (slice(0, 256), slice(8192 - 256, 8192)):
        sub = cols.subset(sl)
        ref, _, _ = ob.run_nomcica(sub)
        g = {k: np.asfortranarray(v[sl]) for k, v in got.items()}
        print("nomcica8192x72", sl, {k: float(np.abs(g[k] - ref[k]).max()) for k in FLUX + HR},
              flush=True)
        assert_close(g, ref, f"overcast8192x72 {sl}")
