"""Several host threads calling the library at once (ctypes releases the GIL for each C call).

* Four threads, each on its own stream, each running the fused McICA forward on its own block of
  columns (different sizes, so each stream's scratch grows differently) and the adjoint: every
  result is bit-for-bit the single-threaded one.
* Four threads sharing one stream, with calls of growing size (the stream's scratch buffer is
  reallocated between calls): calls on one stream run one at a time, and every result is
  bit-for-bit the single-threaded one.
* The host-array entries (default stream) from three threads at once: equal to serial calls.
"""
import threading

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, workload
from test_gpu_parity import sw_args

pytestmark = pytest.mark.gpu

SIZES = (300, 1200, 2500, 4000)


def _run_threads(fns):
    """Run each fn in its own thread; returns their results, re-raising the first error."""
    res, err = [None] * len(fns), [None] * len(fns)

    def work(i):
        try:
            res[i] = fns[i]()
        except BaseException as e:   # noqa: BLE001 -- reported in the main thread
            err[i] = e

    ts = [threading.Thread(target=work, args=(i,)) for i in range(len(fns))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
        assert not t.is_alive(), "a calling thread did not finish"
    for e in err:
        if e is not None:
            raise e
    return res


def _blocks():
    full = workload("mcica8192x72")
    out, c0 = [], 0
    for n in SIZES:
        out.append(full.subset(slice(c0, c0 + n)))
        c0 += n
    return out


def _forward_and_adjoint(api, cols, d, stream):
    import torch
    with torch.cuda.stream(stream):
        f = {}
        for _ in range(3):
            f = api.mcica_rrtmg_sw(cols, d)
        st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
        sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
        g = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh)
    stream.synchronize()
    return {**{k: v.cpu() for k, v in f.items()}, **{"d_" + k: v.cpu() for k, v in g.items()}}


def _serial(api, blocks, ds):
    import torch
    s = torch.cuda.Stream()
    return [_forward_and_adjoint(api, c, d, s) for c, d in zip(blocks, ds)]


def _same(a, b):
    import torch
    for k in a:
        assert torch.equal(a[k], b[k]), k


def test_threads_on_own_streams(gpu_lib):
    import torch
    api = gpu_lib
    blocks = _blocks()
    ds = [api.device_inputs(c) for c in blocks]
    torch.cuda.synchronize()
    ref = _serial(api, blocks, ds)
    streams = [torch.cuda.Stream() for _ in blocks]
    print("threads on own streams", flush=True)
    got = _run_threads([lambda c=c, d=d, s=s: _forward_and_adjoint(api, c, d, s)
                        for c, d, s in zip(blocks, ds, streams)])
    for r, g in zip(ref, got):
        _same(r, g)


def test_threads_sharing_one_stream(gpu_lib):
    import torch
    api = gpu_lib
    blocks = _blocks()
    ds = [api.device_inputs(c) for c in blocks]
    torch.cuda.synchronize()
    ref = _serial(api, blocks, ds)
    shared = torch.cuda.Stream()
    print("threads sharing one stream", flush=True)
    got = _run_threads([lambda c=c, d=d: _forward_and_adjoint(api, c, d, shared)
                        for c, d in zip(blocks, ds)])
    for r, g in zip(ref, got):
        _same(r, g)


def test_threads_host_arrays(gpu_lib):
    api = gpu_lib
    cases = [make_columns(12, 40 + 10 * i, icld=2, iaer=10, seed=70 + i) for i in range(3)]
    mcs = [ob.run_mcica(c)[0] for c in cases]
    ref = [api.rrtmg_sw(**sw_args(c, m)) for c, m in zip(cases, mcs)]
    print("threads on host arrays", flush=True)
    got = _run_threads([lambda c=c, m=m: [api.rrtmg_sw(**sw_args(c, m)) for _ in range(3)][-1]
                        for c, m in zip(cases, mcs)])
    for r, g in zip(ref, got):
        for k in ("swuflx", "swdflx", "swuflxc", "swdflxc", "swhr", "swhrc"):
            assert np.array_equal(r[k], g[k]), k
