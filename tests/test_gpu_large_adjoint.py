"""The large adjoint workload end to end on the GPU: expran65536x137 (65536 columns x 137 layers,
exponential-random overlap, aerosols), the cfg4 case.

* The FP64 table forward on 32 columns sampled over the whole range against the oracle (fluxes
  within 0.1 W/m2, heating rates within 0.01 K/day); the exact-exponential forward against the
  oracle's exact mode (1e-6 W/m2, 1e-4 K/day); the mixed form against the FP64 form on every
  column (0.1 W/m2, 0.01 K/day).
* The adjoint with every layer's state stored and with a checkpoint every 16 layers: the same
  gradients bit for bit, on all 65536 columns.
* Its directional derivative against finite differences of the exact forward on the sampled
  columns (central difference or, where the step crosses a table-cell boundary, a one-sided
  Richardson quotient; each column within 1e-4 relative).
"""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import workload

pytestmark = pytest.mark.gpu

KINDS = ("tlay", "h2ovmr", "o3vmr", "clwp", "ciwp")
ALB = ("asdir", "asdif", "aldir", "aldif")
FLUX = ("swuflx", "swdflx", "swuflxc", "swdflxc")
HR = ("swhr", "swhrc")


@pytest.fixture(scope="module")
def big(gpu_lib):
    cols = workload("expran65536x137")
    d = gpu_lib.device_inputs(cols)
    sample = np.linspace(0, cols.ncol - 1, 32).astype(int)
    yield cols, d, sample
    del d
    import torch
    torch.cuda.empty_cache()


def _host(api, out, sample):
    return {k: api.to_host(v)[sample] for k, v in out.items()}


def test_large_forward_against_oracle(gpu_lib, big):
    import torch
    api = gpu_lib
    cols, d, sample = big
    sub = cols.subset(sample)
    mc, _ = ob.run_mcica(sub)
    for prec, exact, ftol, htol in (("fp64", 0, 0.1, 0.01), ("exact", 1, 1e-6, 1e-4)):
        out = api.mcica_rrtmg_sw(cols, d, precision=prec)
        torch.cuda.synchronize()
        got = _host(api, out, sample)
        del out
        ob.orc().orc_set_exact_exp(exact)
        try:
            ref = ob.run_rrtmg(sub, mc)
        finally:
            ob.orc().orc_set_exact_exp(0)
        e = {k: float(np.abs(got[k] - ref[k]).max()) for k in FLUX + HR}
        print(f"expran65536x137 {prec} forward vs oracle on {len(sample)} columns:", e)
        assert max(e[k] for k in FLUX) <= ftol and max(e[k] for k in HR) <= htol, (prec, e)


def test_large_mixed_against_fp64(gpu_lib, big):
    import torch
    api = gpu_lib
    cols, d, _ = big
    a = api.mcica_rrtmg_sw(cols, d, precision="fp64")
    b = api.mcica_rrtmg_sw(cols, d, precision="mixed")
    torch.cuda.synchronize()
    e = {k: float((a[k] - b[k]).abs().max().item()) for k in a}
    print("expran65536x137 mixed vs fp64, every column:", e)
    assert max(e[k] for k in FLUX) <= 0.1 and max(e[k] for k in HR) <= 0.01, e


def test_large_store_equals_checkpoint(gpu_lib, big):
    import torch
    api = gpu_lib
    cols, d, _ = big
    N, L = cols.ncol, cols.nlay
    st = torch.ones(N, dtype=torch.float64, device="cuda")
    sh = torch.rand(L, N, dtype=torch.float64, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(11))
    a = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=0)
    b = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=16)
    torch.cuda.synchronize()
    for k in a:
        assert torch.isfinite(a[k]).all(), k
        assert torch.equal(a[k], b[k]), k


def test_large_directional_derivative(gpu_lib, big):
    import torch
    from rrtmg_sw_amd import gradcheck
    api = gpu_lib
    cols, _, sample = big
    sub = cols.subset(sample)
    N, L = sub.ncol, sub.nlay
    rng = np.random.default_rng(13)
    dirs = {k: 1e-4 * getattr(sub, k) * rng.uniform(-1, 1, getattr(sub, k).shape)
            for k in KINDS + ALB}
    dsub = api.device_inputs(sub)
    objs = {1: (np.ones(N), np.zeros((N, L))), 2: (np.zeros(N), np.ones((N, L)))}
    for ck in (0, 16):
        def adjoint(st, sh):
            g = api.mcica_rrtmg_sw_adjoint(sub, dsub, api.to_device(st), api.to_device(sh), ckpt=ck)
            torch.cuda.synchronize()
            return {k: api.to_host(v) for k, v in g.items()}

        for name, (rel, n1, n8) in gradcheck.directional(api, sub, dirs, objs, adjoint).items():
            print(f"expran65536x137 J{name} ckpt {ck} directional over {N} columns: relative "
                  f"error median {np.median(rel):.2e}, max {rel.max():.2e}; {n1} one-sided, "
                  f"{n8} at step / 8")
            assert rel.max() <= 1e-4, (name, ck, np.sort(rel)[-5:])
