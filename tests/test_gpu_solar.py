"""Solar-source options through the GCM entry points against the oracle: the day-of-year
Earth-Sun distance (dyofyr > 0, earth_sun, src/rrtmg_sw_rad.f90:844-868), a given solar
constant, and every solar-variability branch of inatm_sw (isolvar -1 to 3 with the optional
bndsolvar, indsolvar and solcycfrac, src/rrtmg_sw_rad.f90:1248-1512), on clear and cloudy McICA
columns (host arrays, and the fused device entry), on the non-McICA driver and through the
adjoint (linear in the incident flux). Bars as in the parity tests; the oracle matches
the compiled reference bit for bit on the same options (test_oracle_reference.py)."""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, workload
from test_gpu_nomcica import nm_args
from test_gpu_parity import assert_close, sw_args

pytestmark = pytest.mark.gpu

SOLVAR = [
    (-1, 0.0, {}), (-1, 1361.0, {"bndsolvar": np.linspace(0.9, 1.1, 14)}),
    (0, 0.0, {}), (1, 1361.0, {"indsolvar": [1.2, 0.8], "solcycfrac": 0.3}),
    (1, 0.0, {"indsolvar": [0.9, 1.1], "solcycfrac": 0.7}), (2, 0.0, {"indsolvar": [0.16, 1200.0]}),
    (3, 1361.0, {"bndsolvar": np.linspace(1.05, 0.95, 14)}), (3, 0.0, {})]


@pytest.mark.parametrize("isolvar,scon,opts", SOLVAR)
def test_solar_options_mcica_host(gpu_lib, isolvar, scon, opts):
    api = gpu_lib
    cols = make_columns(8, 40, icld=2, iaer=10, seed=51)
    cols.isolvar, cols.scon, cols.dyofyr = isolvar, scon, 172
    print(f"isolvar={isolvar} scon={scon} {sorted(opts)}", flush=True)
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc, **opts)
    kw = {k: (np.asarray(v, dtype=np.float64) if k != "solcycfrac" else v) for k, v in opts.items()}
    got = api.rrtmg_sw(**sw_args(cols, mc), **kw)
    assert_close(got, ref, f"isolvar={isolvar}")


def test_solar_options_clear_and_nomcica(gpu_lib):
    api = gpu_lib
    cols = workload("clear64x60", ncol=6)
    cols.isolvar, cols.scon, cols.dyofyr = 1, 1361.0, 31
    opts = {"indsolvar": np.array([1.2, 0.8]), "solcycfrac": 0.3}
    ref = ob.run_rrtmg(cols, **opts)
    got = api.rrtmg_sw(**sw_args(cols, ob.empty_mcica(cols)), **opts)
    assert_close(got, ref, "clear isolvar=1")
    oc = workload("overcast8192x72", ncol=4)
    oc.isolvar, oc.scon = 3, 1361.0
    sv = (np.linspace(0.9, 1.1, 14), np.array([1.2, 0.8]), 0.3)
    nref, _, _ = ob.run_nomcica(oc, solvar=sv)
    got = api.rrtmg_sw_nomcica(**nm_args(oc), bndsolvar=sv[0], indsolvar=sv[1], solcycfrac=sv[2])
    assert_close(got, nref, "nomcica isolvar=3")


def test_solar_options_fused_and_adjoint(gpu_lib):
    """The fused device entry takes the same optional arguments (host arrays) as rrtmg_sw, and the
    adjoint sees them: scaling every band's source by 2 (isolvar = -1, bndsolvar = 2) doubles
    every gradient, as the path is linear in the incident flux."""
    import torch
    api = gpu_lib
    cols = make_columns(16, 40, icld=5, iaer=10, seed=53)
    cols.dyofyr = 100
    d = api.device_inputs(cols)
    mc, _ = ob.run_mcica(cols)
    for isolvar, scon, opts in SOLVAR:
        cols.isolvar, cols.scon = isolvar, scon
        ref = ob.run_rrtmg(cols, mc, **opts)
        out = api.mcica_rrtmg_sw(cols, d, **opts)
        torch.cuda.synchronize()
        assert_close({k: api.to_host(v) for k, v in out.items()}, ref, f"fused isolvar={isolvar}")
    cols.isolvar, cols.scon = -1, 0.0
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
    g1 = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, bndsolvar=np.ones(14))
    g2 = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, bndsolvar=np.full(14, 2.0))
    torch.cuda.synchronize()
    for k in g1:
        assert g1[k].abs().max() > 0, k
        assert torch.allclose(g2[k], 2.0 * g1[k], rtol=1e-13, atol=0.0), k
