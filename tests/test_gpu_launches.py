"""Launch-level properties of the HIP path at full width.

* Columns are independent and their results do not depend on how a call is cut into launches:
  with RRTMGSW_MAX_LAUNCH_COLUMNS lowering the columns per launch (so that the per-launch column
  offset, the XCD-aware order within a partial last launch and the reuse of the scratch buffer
  are all exercised), every output of the fused McICA forward (FP64, mixed, exact-exponential),
  the non-McICA forward and the adjoint (store and checkpoint-16) is bit-for-bit the single-launch
  one.
* The full 8192 x 72 McICA workload in one call, checked against the oracle on 48 columns drawn
  from the whole range (per-column KISS streams make a column subset exact to compare), and the
  level-wise properties every column must satisfy: the downward flux at the top is the same
  incident flux for both skies, fluxes are non-negative, the upward flux at the top does not
  exceed the downward one.
"""
import os

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload
from test_gpu_nomcica import nm_args
from test_gpu_parity import assert_close

pytestmark = pytest.mark.gpu

KEY = "RRTMGSW_MAX_LAUNCH_COLUMNS"


def _chunked(fn, ncols):
    os.environ[KEY] = str(ncols)
    try:
        return fn()
    finally:
        del os.environ[KEY]


def _same(a, b):
    import torch
    torch.cuda.synchronize()
    for k in a:
        assert torch.equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed", "exact"])
def test_chunked_forward_is_bit_identical(gpu_lib, precision):
    from rrtmg_sw_amd import api
    cols = workload("mcica8192x72", ncol=1000)
    d = api.device_inputs(cols)
    one = api.mcica_rrtmg_sw(cols, d, precision=precision)
    many = _chunked(lambda: api.mcica_rrtmg_sw(cols, d, precision=precision), 96)
    _same(one, many)


def test_chunked_nomcica_is_bit_identical(gpu_lib):
    from rrtmg_sw_amd import api
    cols = workload("overcast8192x72", ncol=700)
    d = api.device_inputs(cols)
    kw = {**nm_args(cols), **{k: d[n] for k, n in (
        ("play", "play"), ("plev", "plev"), ("tlay", "tlay"), ("h2ovmr", "h2ovmr"), ("o3vmr", "o3vmr"),
        ("co2vmr", "co2vmr"), ("ch4vmr", "ch4vmr"), ("o2vmr", "o2vmr"), ("asdir", "asdir"),
        ("asdif", "asdif"), ("aldir", "aldir"), ("aldif", "aldif"), ("coszen", "coszen"),
        ("cldfr", "cldfrac"), ("taucld", "tauc"), ("ssacld", "ssac"), ("asmcld", "asmc"),
        ("fsfcld", "fsfc"), ("cicewp", "ciwp"), ("cliqwp", "clwp"), ("reice", "rei"),
        ("reliq", "rel"))}, "tlev": None, "tsfc": None, "n2ovmr": None, "tauaer": None,
        "ssaaer": None, "asmaer": None, "ecaer": None}
    one = api.rrtmg_sw_nomcica(**kw)
    many = _chunked(lambda: api.rrtmg_sw_nomcica(**kw), 100)
    _same({k: v for k, v in one.items() if k not in ("icld", "iaer")},
          {k: v for k, v in many.items() if k not in ("icld", "iaer")})


@pytest.mark.parametrize("ckpt", [0, 16])
def test_chunked_adjoint_is_bit_identical(gpu_lib, ckpt):
    import torch
    from rrtmg_sw_amd import api
    cols = workload("expran4096x60", ncol=600)
    d = api.device_inputs(cols)
    g = torch.Generator(device="cpu").manual_seed(5)
    st = torch.rand(cols.ncol, generator=g, dtype=torch.float64).cuda()
    sh = torch.rand(cols.nlay, cols.ncol, generator=g, dtype=torch.float64).cuda()
    one = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ckpt)
    many = _chunked(lambda: api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ckpt), 70)
    _same(one, many)


def test_full_width_sample_against_oracle(gpu_lib):
    import torch
    from rrtmg_sw_amd import api
    cols = workload("mcica8192x72")
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d)
    torch.cuda.synchronize()
    got = {k: api.to_host(v) for k, v in out.items()}
    L = cols.nlay
    # every column: same incident flux at the top for both skies, non-negative fluxes
    assert np.array_equal(got["swdflx"][:, L], got["swdflxc"][:, L])
    for k in ("swuflx", "swdflx", "swuflxc", "swdflxc"):
        assert (got[k] >= 0.0).all(), k
    assert (got["swuflx"][:, L] <= got["swdflx"][:, L]).all()
    idx = np.sort(np.random.default_rng(7).choice(cols.ncol, 48, replace=False))
    sub = cols.subset(idx)
    mc, _ = ob.run_mcica(sub)
    ref = ob.run_rrtmg(sub, mc)
    g = {k: np.asfortranarray(v[idx]) for k, v in got.items()}
    print("full-width sample", {k: float(np.abs(g[k] - ref[k]).max()) for k in ref if k.startswith("sw")},
          flush=True)
    assert_close(g, ref, "mcica8192x72 sample")
