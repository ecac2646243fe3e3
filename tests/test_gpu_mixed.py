"""Mixed FP32/FP64 form of the fused McICA path against the FP64 path and the oracle.

Bars (written here): every flux (up/down, total and clear sky, at every level) within
0.1 W/m2 and every heating rate within 0.01 K/day of the FP64 reference. The sub-column masks
are the same integer streams in both forms. Each test prints its per-output maximum errors.
"""
import json

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload

pytestmark = pytest.mark.gpu

FLUX = ("swuflx", "swdflx", "swuflxc", "swdflxc")
HR = ("swhr", "swhrc")
FLUX_TOL = 0.1     # W/m2
HR_TOL = 0.01      # K/day


def _run(api, cols, precision):
    import torch
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d, precision=precision)
    torch.cuda.synchronize()
    return {k: api.to_host(v) for k, v in out.items()}


def _errors(got, ref):
    e = {k: float(np.abs(got[k] - ref[k]).max()) for k in FLUX + HR}
    # heating-rate error by layer (max over columns), top layers last
    e["swhr_by_layer_top8"] = [float(x) for x in np.abs(got["swhr"] - ref["swhr"]).max(axis=0)[-8:]]
    return e


def _check(e, what):
    for k in FLUX:
        assert e[k] <= FLUX_TOL, (what, k, e[k])
    for k in HR:
        assert e[k] <= HR_TOL, (what, k, e[k])


@pytest.mark.parametrize("name", ["mcica8192x72", "expran4096x60"])
def test_mixed_vs_fp64_full_workload(gpu_lib, name):
    cols = workload(name)
    ref = _run(gpu_lib, cols, "fp64")
    got = _run(gpu_lib, cols, "mixed")
    e = _errors(got, ref)
    print(f"mixed vs fp64 {name}:", json.dumps(e))
    _check(e, name)


def test_mixed_vs_oracle_slice(gpu_lib):
    cols = workload("mcica8192x72", ncol=96)
    got = _run(gpu_lib, cols, "mixed")
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc)
    e = _errors(got, ref)
    print("mixed vs oracle:", json.dumps(e))
    _check(e, "oracle")
