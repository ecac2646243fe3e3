"""Seeded random cases mixing the options the single-purpose tests vary one at a time: column
and layer counts (1-24 columns, 4-100 layers), overlap (icld 0-5), generator (KISS / Mersenne
Twister), aerosols (iaer 0 / 6 / 10, confined to the lower troposphere as the workloads' are:
an aerosol optical depth of 0.1 in a 0.002 hPa layer heats it by 1e5 K/day, where the mixed
form's FP32 relative precision alone exceeds the absolute heating-rate bar), ice optics (iceflag 1 / 2 / 3), solar variability
(isolvar -1 / 0) and sun angles. Each case runs the host-array pair (mcica_subcol_sw +
rrtmg_sw), the fused device entry (FP64 and mixed) and checks them against the oracle: McICA
arrays bit-exact, fluxes within 0.1 W/m2 and heating rates within 0.01 K/day."""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns
from test_gpu_parity import FLUX, HR, assert_close, mc_args, sw_args

pytestmark = pytest.mark.gpu

NCASES = 64


def _case(seed, max_nlay=100, min_icld=0):
    rng = np.random.default_rng(7000 + seed)
    ncol = int(rng.integers(1, 25))
    nlay = int(rng.integers(4, max_nlay + 1))
    icld = int(rng.integers(min_icld, 6))
    iaer = int(rng.choice([0, 6, 10]))
    cols = make_columns(ncol, nlay, icld=icld, iaer=iaer, seed=7100 + seed)
    if iaer == 6:   # ECMWF aerosol amounts in the layers below 700 hPa (about 3 km)
        ec = rng.uniform(0.0, 0.2, cols.ecaer.shape)
        ec[cols.play < 700.0] = 0.0
        cols.ecaer = np.asfortranarray(ec)
    cols.irng = int(rng.integers(0, 2))
    cols.permuteseed = int(rng.integers(0, 5000))
    cols.iceflgsw = int(rng.choice([1, 2, 3]))   # the drawn ice sizes (20-120 um) suit all three
    cols.isolvar = int(rng.choice([-1, 0]))
    cols.coszen = np.asfortranarray(rng.uniform(0.02, 1.0, cols.coszen.shape))
    return cols


@pytest.mark.parametrize("seed", range(NCASES))
def test_random_case(gpu_lib, seed):
    import torch
    from rrtmg_sw_amd import api
    cols = _case(seed)
    label = (f"seed={seed} ncol={cols.ncol} nlay={cols.nlay} icld={cols.icld} irng={cols.irng} "
             f"iaer={cols.iaer} iceflag={cols.iceflgsw} isolvar={cols.isolvar}")
    print(label, flush=True)
    ref_mc, ref_irng = ob.run_mcica(cols) if cols.icld else (ob.empty_mcica(cols), cols.irng)
    ref = ob.run_rrtmg(cols, ref_mc)
    # host-array pair, as a GCM calls the reference
    if cols.icld:
        mc = gpu_lib.mcica_subcol_sw(**mc_args(cols))
        assert mc["irng"] == ref_irng, label
        for k in ref_mc:
            assert np.array_equal(mc[k], ref_mc[k]), (label, k)
    got = gpu_lib.rrtmg_sw(**sw_args(cols, ref_mc))
    assert_close(got, ref, label + " host")
    # fused device entry, FP64 and mixed
    if cols.icld:
        d = api.device_inputs(cols)
        for precision in ("fp64", "mixed"):
            out = api.mcica_rrtmg_sw(cols, d, precision=precision)
            torch.cuda.synchronize()
            fused = {k: api.to_host(v) for k, v in out.items()}
            if precision == "fp64":
                assert_close(fused, ref, f"{label} fused fp64")
            for k in FLUX + HR:   # the bars of the precision form
                dmax = float(np.abs(fused[k] - ref[k]).max())
                assert dmax <= (0.01 if k in HR else 0.1), (label, precision, k, dmax)
    print("  max |d|", {k: float(np.abs(np.asarray(got[k]) - ref[k]).max()) for k in FLUX + HR})


@pytest.mark.parametrize("seed", range(32))
def test_random_nomcica_case(gpu_lib, seed):
    """The same draws through the non-McICA driver, every cloudy layer made overcast (the driver
    stops on partial clouds)."""
    from rrtmg_sw_amd.workloads import overcast
    from test_gpu_nomcica import nm_args
    cols = overcast(_case(seed))
    label = (f"seed={seed} ncol={cols.ncol} nlay={cols.nlay} icld={cols.icld} iaer={cols.iaer} "
             f"iceflag={cols.iceflgsw} isolvar={cols.isolvar}")
    print(label, flush=True)
    got = gpu_lib.rrtmg_sw_nomcica(**nm_args(cols))
    ref, ic, ia = ob.run_nomcica(cols)
    assert (got["icld"], got["iaer"]) == (ic, ia), label
    assert_close(got, ref, label + " nomcica")


@pytest.mark.parametrize("seed", range(8))
def test_random_adjoint_case(gpu_lib, seed):
    """The same draws (at most 60 layers) through the adjoint: store and checkpoint-16 (or nlay)
    give the same bits, and the directional derivatives of J1 = TOA upward flux and J2 = a
    weighted sum of the heating rates agree with finite differences of the exact-exponential forward within 1e-4
    relative in every column (rrtmg_sw_amd.gradcheck)."""
    import torch
    from rrtmg_sw_amd import gradcheck
    from test_gpu_edges import ALB, KINDS, _adjoint
    api = gpu_lib
    cols = _case(100 + seed, max_nlay=60, min_icld=1)
    N, L = cols.ncol, cols.nlay
    print(f"seed={seed} ncol={N} nlay={L} icld={cols.icld} iaer={cols.iaer} irng={cols.irng}", flush=True)
    d = api.device_inputs(cols)
    rng = np.random.default_rng(7400 + seed)
    st = torch.ones(N, dtype=torch.float64, device="cuda")
    sh = torch.zeros(L, N, dtype=torch.float64, device="cuda")
    ref = _adjoint(api, cols, d, st, sh, 0)
    g16 = _adjoint(api, cols, d, st, sh, min(16, L))
    for name in ref:
        assert torch.equal(ref[name], g16[name]), name
    pdp = cols.plev[:, :-1] - cols.plev[:, 1:]
    w = rng.uniform(0.0, 1.0, (N, L)) * pdp / pdp.max()
    dirs = {k: 1e-4 * getattr(cols, k) * rng.uniform(-1, 1, getattr(cols, k).shape)
            for k in KINDS + ALB}
    objs = {1: (np.ones(N), np.zeros((N, L))), 2: (np.zeros(N), w)}

    def adjoint(s1, s2):
        g = _adjoint(api, cols, d, api.to_device(s1), api.to_device(s2), 0)
        return {k: api.to_host(v) for k, v in g.items()}

    for obj, (rel, n1, n8) in gradcheck.directional(api, cols, dirs, objs, adjoint).items():
        print(f"  J{obj}: max relative error {rel.max():.2e}; {n1} one-sided, {n8} at step / 8")
        assert rel.max() <= 1e-4, (seed, obj, rel)
