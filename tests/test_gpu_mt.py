"""The Mersenne Twister stream of a long call is generated in segments of 2^19 words (or more, at
most 256 segments), one workgroup each, from start states formed all at once by jump-ahead
(x^(k J) mod the MT19937 characteristic polynomial for every segment k, on the host). Checks: the McICA sub-column arrays of calls whose streams span
several segments bit-for-bit equal to the oracle's (which draws the stream one word at a time),
for icld 2 (one draw per cell) and 5 (two); and the full 8192 x 72 fused call identical to the
one-workgroup generator (RRTMGSW_MT_SEQUENTIAL=1)."""
import os

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, workload
from test_gpu_parity import mc_args

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ncol,icld", [(600, 2), (330, 5)])
def test_jumped_stream_masks_match_oracle(gpu_lib, ncol, icld):
    cols = make_columns(ncol, 72, icld=icld, seed=17 + icld)
    cols.irng = 1
    draws = 112 * ncol * 72 * (2 if icld >= 4 else 1)
    assert draws > 2 * 2 ** 19   # more than two segments: the jump-ahead path
    print(f"ncol={ncol} icld={icld} draws={draws}", flush=True)
    got = gpu_lib.mcica_subcol_sw(**mc_args(cols))
    ref, _ = ob.run_mcica(cols)
    for k in ref:
        assert np.array_equal(got[k], ref[k]), k


def test_jumped_stream_equals_sequential_full_width(gpu_lib):
    import torch
    api = gpu_lib
    cols = workload("mcica8192x72")
    cols.irng = 1
    d = api.device_inputs(cols)
    jumped = api.mcica_rrtmg_sw(cols, d)
    torch.cuda.synchronize()
    os.environ["RRTMGSW_MT_SEQUENTIAL"] = "1"
    try:
        seq = api.mcica_rrtmg_sw(cols, d)
        torch.cuda.synchronize()
    finally:
        del os.environ["RRTMGSW_MT_SEQUENTIAL"]
    for k in jumped:
        assert torch.equal(jumped[k], seq[k]), k
