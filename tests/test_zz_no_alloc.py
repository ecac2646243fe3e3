"""rrtmgsw_mcica_sw_dev_noalloc (the entry point that allocates no device memory) against
rrtmgsw_mcica_sw_dev, through the C-ABI: the same kernels per column, so every output must be
the same bits. Two calls in a row, the second with a smaller input (fewer columns, other layers,
overlap and aerosols), so the second runs in an array that still holds the first call's data;
every precision form; the Mersenne Twister generator; a stream too large for the array refused
with a clear error.
"""
import numpy as np
import pytest

from rrtmg_sw_amd.workloads import workload

pytestmark = pytest.mark.gpu


def _pair(api, cols, precision="fp64"):
    import torch
    d = api.device_inputs(cols)
    ref = api.mcica_rrtmg_sw(cols, d, precision=precision)
    got = api.mcica_rrtmg_sw(cols, d, precision=precision, noalloc=True)
    torch.cuda.synchronize()
    return ref, got


def _same(ref, got, what):
    import torch
    for k in ref:
        assert torch.equal(ref[k], got[k]), (what, k, float((ref[k] - got[k]).abs().max().item()))


def test_noalloc_matches_allocating_entry(gpu_lib):
    api = gpu_lib
    big = workload("mcica8192x72")                      # 14 launches of 629 columns
    small = workload("expran4096x60", ncol=300)         # one launch, icld 5 + aerosols
    for prec in ("fp64", "mixed", "exact"):
        ref, got = _pair(api, big, prec)
        _same(ref, got, ("mcica8192x72", prec))
        ref, got = _pair(api, small, prec)
        _same(ref, got, ("expran4096x60[:300]", prec))


def test_noalloc_mersenne_twister(gpu_lib):
    api = gpu_lib
    cols = workload("mcica8192x72", ncol=96)
    cols.irng = 1
    ref, got = _pair(api, cols)
    _same(ref, got, "irng = 1")


def test_noalloc_refuses_a_stream_too_large(gpu_lib):
    """icld 4/5 with the Mersenne Twister draws 2 x 112 x ncol x nlay words: 504 MiB for
    8192 x 72, more than the 256 MiB array."""
    import torch
    api = gpu_lib
    cols = workload("mcica8192x72")
    cols.icld, cols.irng = 5, 1
    d = api.device_inputs(cols)
    with pytest.raises(api.RRTMGSWError) as e:
        api.mcica_rrtmg_sw(cols, d, noalloc=True)
    assert e.value.code == 2 and "does not fit" in e.value.message, e.value
    # the array is still usable afterwards
    ref, got = _pair(api, workload("expran4096x60", ncol=64))
    _same(ref, got, "after a refused call")
    torch.cuda.synchronize()
