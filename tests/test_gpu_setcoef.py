"""setcoef_sw on the GPU (rrtmgsw_setcoef_dev, the forward kernels' own setcoef code) against the
oracle's setcoef_sw, fed from the same inatm_sw layer amounts.

Bar (written here): laytrop, laylow, laysolfr, jp, jt, jt1, indself, indfor bit-exact; the column
amounts, selffac, selffrac, forfac and forfrac bit-exact (the same IEEE operations in the same
order); fac00..fac11 within 4e-14 absolute: they hold fp = 5 (preflog(jp) - log p), and the
device log may differ from the host's by an ulp (8.9e-16 at log p = 7).
Cases: the full 8192 x 72 McICA workload, the 4096 x 60 adjoint workload, the extreme columns,
and a grid with the 32 doubles on either side of every pressure switch (jp, laytrop, laylow)
and every temperature switch (jt, jt1 in each of the 58 pressure cells; indfor, indself).
"""
import numpy as np
import pytest

import oracle_bind as ob
import setcoef_grid as sg
from rrtmg_sw_amd.workloads import extreme, workload

pytestmark = pytest.mark.gpu

FAC_TOL = 4e-14
INTS = ("laytrop", "laylow", "jp", "jt", "jt1", "indself", "indfor")
EXACT = ("colh2o", "colco2", "colo3", "colch4", "colo2", "colmol", "selffac", "selffrac",
         "forfac", "forfrac")
FACS = ("fac00", "fac01", "fac10", "fac11")
PROF = ("play", "plev", "tlay", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "o2vmr")


def _compare(api, prof, what):
    import torch
    n, L = prof["play"].shape
    d = {k: api.to_device(prof[k]) for k in PROF}
    got = api.setcoef_sw(n, L, *[d[k] for k in PROF])
    torch.cuda.synchronize()
    g = {k: v.cpu().numpy().T for k, v in got.items()}   # back to Fortran (ncol, ...) shapes
    ref = ob.run_setcoef(n, L, *[prof[k] for k in PROF])
    for k in INTS:
        assert np.array_equal(g[k], ref[k]), (what, k, int((g[k] != ref[k]).sum()))
    # laysolfr: the lower-atmosphere bands have none without a troposphere layer
    lower = [2, 3, 4, 5, 6, 7, 8, 9, 10]   # bands 18-26
    ok = ref["laytrop"] > 0
    for b in range(14):
        sel = ok if b in lower else slice(None)
        assert np.array_equal(g["laysolfr"][sel, b], ref["laysolfr"][sel, b]), (what, "laysolfr", b)
    for k in EXACT:
        assert np.array_equal(g[k], ref[k]), (what, k, float(np.abs(g[k] - ref[k]).max()))
    worst = max(float(np.abs(g[k] - ref[k]).max()) for k in FACS)
    same = np.mean([np.mean(g[k] == ref[k]) for k in FACS])
    print(f"{what}: {n} x {L}: indices and amounts bit-exact; fac00..fac11 max |diff| "
          f"{worst:.1e}, {100 * same:.2f} % bit-identical")
    assert worst <= FAC_TOL, (what, worst)
    return g, ref


def _prof(cols):
    return {k: getattr(cols, k) for k in PROF}


@pytest.mark.parametrize("name", ["mcica8192x72", "expran4096x60"])
def test_setcoef_full_workload(gpu_lib, name):
    _compare(gpu_lib, _prof(workload(name)), name)


def test_setcoef_extreme(gpu_lib):
    _compare(gpu_lib, _prof(extreme()), "extreme")


def test_setcoef_pressure_switches(gpu_lib):
    sw = sg.pressure_switches()
    g, ref = _compare(gpu_lib, sg.pressure_grid(sw), "pressure switches")
    # the grid straddles each switch: both sides occur in every column
    side = np.arange(-sg.NEAR, sg.NEAR) >= 0
    names = list(sw)
    for i, name in enumerate(names):
        if isinstance(name, tuple):
            assert np.array_equal(g["jp"][i] < name[1], side), name
    assert g["laytrop"][names.index("trop")] == sg.NEAR
    assert g["laylow"][names.index("low")] == sg.NEAR


def test_setcoef_temperature_switches(gpu_lib):
    _compare(gpu_lib, sg.temperature_grid(), "temperature switches")
