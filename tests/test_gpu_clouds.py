"""The McICA drivers and the adjoint under every cloud-optics option of cldprmc_sw
(src/rrtmg_sw_cldprmc.f90): inflag 2 with the Ebert-Curry (iceflag 1), Key/Streamer (2) and Fu
(3) ice parameterizations, and inflag 0 (per-band optical properties given by the caller).

* Forward: the fused device entry and rrtmg_sw on host arrays against the oracle (fluxes within
  0.1 W/m2, heating rates within 0.01 K/day, median flux difference at the rounding level).
* Adjoint: directional derivatives of J1 = TOA upward flux and J2 = mass-weighted heating rates
  against central differences of the exact-exponential forward (step 1e-4 |x| on every input at
  once; 90th percentile of the columns' relative errors within 1e-4, all within 1e-3). With
  inflag 0 the water paths do not reach the optics: their gradients are exactly zero.
"""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, with_cloud_optics
from test_gpu_adjoint import ALB, KINDS
from test_gpu_edges import _adjoint, _err, _fused, _toa_and_hr
from test_gpu_parity import assert_close, sw_args

pytestmark = pytest.mark.gpu

OPTIONS = [(2, 1), (2, 2), (2, 3), (0, 3)]


def _cols(inflag, iceflag, ncol=16, seed=41):
    return with_cloud_optics(make_columns(ncol, 40, icld=2, iaer=10, seed=seed), inflag, iceflag,
                             seed=seed)


@pytest.mark.parametrize("inflag,iceflag", OPTIONS)
def test_cloud_options_forward(gpu_lib, inflag, iceflag):
    api = gpu_lib
    cols = _cols(inflag, iceflag)
    print(f"inflag={inflag} iceflag={iceflag}", flush=True)
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc)
    assert np.abs(ref["swuflx"] - ref["swuflxc"]).max() > 1.0     # the clouds matter
    got = _fused(api, cols)
    print("fused", _err(got, ref))
    assert_close(got, ref, f"fused {inflag}/{iceflag}")
    got = api.rrtmg_sw(**sw_args(cols, mc))
    assert_close(got, ref, f"host {inflag}/{iceflag}")


@pytest.mark.parametrize("inflag,iceflag", OPTIONS)
def test_cloud_options_adjoint(gpu_lib, inflag, iceflag):
    import torch
    api = gpu_lib
    cols = _cols(inflag, iceflag)
    N, L = cols.ncol, cols.nlay
    print(f"adjoint inflag={inflag} iceflag={iceflag}", flush=True)
    d = api.device_inputs(cols)
    rng = np.random.default_rng(7)
    pdp = cols.plev[:, :-1] - cols.plev[:, 1:]
    w = rng.uniform(0.0, 1.0, (N, L)) * pdp / pdp.max()
    dirs = {k: 1e-4 * getattr(cols, k) * rng.uniform(-1, 1, getattr(cols, k).shape)
            for k in KINDS + ALB}
    plus, minus = cols.subset(np.arange(N)), cols.subset(np.arange(N))
    for k, dx in dirs.items():
        getattr(plus, k)[...] += dx
        getattr(minus, k)[...] -= dx
    up_p, hr_p = _toa_and_hr(api, plus)
    up_m, hr_m = _toa_and_hr(api, minus)
    one = torch.ones(N, dtype=torch.float64, device="cuda")
    zero_hr = torch.zeros(L, N, dtype=torch.float64, device="cuda")
    w_dev = torch.from_numpy(np.ascontiguousarray(w.T)).cuda()
    for obj, (s1, s2) in ((1, (one, zero_hr)), (2, (torch.zeros_like(one), w_dev))):
        g = {k: api.to_host(v) for k, v in _adjoint(api, cols, d, s1, s2, 0).items()}
        if inflag == 0:
            assert not g["clwp"].any() and not g["ciwp"].any()
        lin = sum((g[k] * dirs[k]).sum(axis=1) for k in KINDS)
        lin = lin + sum(g["alb"][:, m] * dirs[k] for m, k in enumerate(ALB))
        fdv = (up_p - up_m) / 2 if obj == 1 else ((hr_p - hr_m) * w).sum(axis=1) / 2
        rel = np.abs(lin - fdv) / np.abs(fdv)
        print(f"J{obj}: sorted relative errors " + " ".join(f"{x:.1e}" for x in np.sort(rel)))
        assert np.quantile(rel, 0.9) <= 1e-4 and rel.max() <= 1e-3, (obj, rel)
