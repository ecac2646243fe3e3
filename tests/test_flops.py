"""The operation-counting build of the oracle (oracle/orc_count.cpp), which gives the algorithmic
FLOP figures of rrtmg_sw_amd/data/flops_counted.json (written by oracle/count_flops.py):

* it computes what the plain oracle computes, bit for bit (the counted type changes no value);
* its count of vrtqdr_sw is the reference's: 14 operations per layer in the upward pass, 15 per
  layer below the top in the downward pass and 14 per level for the fluxes
  (src/rrtmg_sw_vrtqdr.f90:120-169), 43 klev - 1 in all;
* counts add over column chunks (up to a few operations per call), and the stored 64-column
  samples recount exactly;
* the whole-workload count agrees within 10 % with the routine-by-routine tally of
  rrtmg_sw_amd/flops.py, an independent reading of the reference's sources.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd import flops
from rrtmg_sw_amd.workloads import workload

pytestmark = pytest.mark.skipif(not os.path.exists(ob.COUNT_PATH),
                                reason="oracle/liborc_count.so not built")
TABLE = os.path.join(os.path.dirname(flops.__file__), "data", "flops_counted.json")


def _counted(fn):
    ob.orc_count().orc_fp_reset()
    fn()
    return ob.fp_counts()


def test_counting_build_matches_oracle():
    cols = workload("expran4096x60", ncol=24)
    mc, _ = ob.run_mcica(cols)
    mc2, _ = ob.run_mcica(cols, "count")
    for k in mc:
        assert np.array_equal(mc[k], mc2[k]), k
    a = ob.run_rrtmg(cols, mc)
    b = ob.run_rrtmg(cols, mc2, which="count")
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert np.array_equal(a[k], b[k]), k
    nm = workload("overcast8192x72", ncol=16)
    a, _, _ = ob.run_nomcica(nm)
    b, _, _ = ob.run_nomcica(nm, which="count")
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("klev", [4, 60, 137])
def test_vrtqdr_count(klev):
    rng = np.random.default_rng(klev)
    n = klev + 1
    ins = [rng.uniform(0.0, 0.5, n) for _ in range(5)]
    ptdbt = np.sort(rng.uniform(0.1, 1.0, n))[::-1].copy()
    outs = [np.zeros(n) for _ in range(5)]
    p = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in ins + outs + [ptdbt]]
    lib = ob.orc_count()
    ops, trans = _counted(lambda: lib.orc_vrtqdr(C.c_int(klev), p[0], p[1], p[2], p[3], p[4],
                                                 p[5], p[6], p[7], p[10], p[8], p[9]))
    assert (ops, trans) == (43 * klev - 1, 0)


def test_counts_add_over_chunks_and_samples_recount():
    with open(TABLE) as f:
        table = json.load(f)
    cols = workload("mcica8192x72", ncol=64)

    def both(c):
        mc, _ = ob.run_mcica(c, "count")
        ob.run_rrtmg(c, mc, which="count")

    whole = _counted(lambda: both(cols))
    halves = [_counted(lambda: both(cols.subset(slice(a, a + 32)))) for a in (0, 32)]
    quarters = [_counted(lambda: both(cols.subset(slice(a, a + 16)))) for a in range(0, 64, 16)]
    # per column, plus a few operations per call (the solar scalars of inatm_sw)
    per_call = sum(h[0] for h in halves) - whole[0]
    assert 0 <= per_call <= 16
    assert sum(q[0] for q in quarters) - whole[0] == 3 * per_call
    assert whole[1] == sum(q[1] for q in quarters)
    rec = table["mcica8192x72"]["first64"]
    assert whole[0] == rec["mcica_subcol_sw"] + rec["rrtmg_sw"]
    assert whole[1] == rec["exp_log_pow"]


@pytest.mark.parametrize("name", ["mcica8192x72", "expran4096x60"])
def test_count_agrees_with_tally(name):
    with open(TABLE) as f:
        rec = json.load(f)[name]
    tally = flops.call_flops(rec["ncol"], rec["nlay"], rec["cloudy_cells"])
    assert abs(rec["total"] / tally - 1.0) <= 0.10, (rec["total"], tally)
    assert flops.workload_flops(name) == rec["total"]
