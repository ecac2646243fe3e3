"""The HIP library at the edges of its input domain, against the C oracle.

* The fewest (4) and most (256) layers the library takes, on every entry that runs the column
  kernels: the fused McICA forward (FP64 table form, and the mixed form against it), rrtmg_sw on
  host arrays with the oracle's sub-columns, and the adjoint (checkpointing at K = 1, 7, 16 and
  K = nlay against storing every layer, and a directional derivative against central
  differences of the exact-exponential forward). 256 layers fill eight 32-layer mask words
  exactly and take the adjoint workgroup past 64 KB of LDS. (The non-McICA driver at 4 and 256
  layers runs among the oracle cases of test_gpu_nomcica.)
* One column (a grid of eight workgroups, seven without a column) and no columns (every entry
  returns at once); rrtmgsw_finalize then rrtmgsw_init (same results after).
* Night and grazing sun: coszen from -0.5 to 1e-12, and extreme atmospheres (temperatures off
  the reference grid, no water vapour or ozone, a 0.001 hPa top), against the oracle on every
  forward driver, the adjoint finite; the mixed form against FP64 and the exact-exponential form
  finite on both. The adjoint under random, maximum and exponential overlap and the Mersenne
  Twister generator (store = checkpoint-7, TOA-flux directional derivative).
* Layer counts outside 4..256 and the reference's STOP conditions, each with its error code
  and message: invalid icld (src/mcica_subcol_gen_sw.f90:262-263), inflag = 1 with McICA
  (src/rrtmg_sw_cldprmc.f90:166), pressures not ordered bottom-up for the KISS seed
  (src/mcica_subcol_gen_sw.f90:470-475) and particle sizes outside the table bounds of each
  parameterization (src/rrtmg_sw_cldprmc.f90:183-184, :213, :236, :273-274;
  src/rrtmg_sw_cldprop.f90:194, :226, :250, :290), with the oracle raising the same message.
Bars as in the parity tests: fluxes within 0.1 W/m2 and heating rates within 0.01 K/day of
the oracle with the median flux difference at the rounding level; masks bit-exact.
"""
import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, overcast
from test_gpu_adjoint import ALB, KINDS
from test_gpu_parity import FLUX, HR, assert_close, mc_args, sw_args
from test_gpu_nomcica import nm_args

pytestmark = pytest.mark.gpu

# (nlay, icld, seed): seeds whose columns hold clouds at both layer counts
LIMITS = [(4, 2, 11), (4, 5, 14), (256, 2, 47), (256, 5, 48)]


def _cols(nlay, icld, seed, ncol=8):
    return make_columns(ncol, nlay, icld=icld, iaer=10, seed=seed)


def _fused(api, cols, precision="fp64"):
    import torch
    out = api.mcica_rrtmg_sw(cols, api.device_inputs(cols), precision=precision)
    torch.cuda.synchronize()
    return {k: api.to_host(v) for k, v in out.items()}


def _err(got, ref):
    return {k: float(np.abs(got[k] - ref[k]).max()) for k in FLUX + HR}


@pytest.mark.parametrize("nlay,icld,seed", LIMITS)
def test_fused_layer_limits(gpu_lib, nlay, icld, seed):
    cols = _cols(nlay, icld, seed)
    print(f"fused nlay={nlay} icld={icld} seed={seed}", flush=True)
    got = _fused(gpu_lib, cols)
    mc, _ = ob.run_mcica(cols)
    assert mc["cldfmcl"].any()
    ref = ob.run_rrtmg(cols, mc)
    print(_err(got, ref))
    assert_close(got, ref, f"nlay={nlay}")
    mixed = _fused(gpu_lib, cols, "mixed")
    e = _err(mixed, got)
    print("mixed vs fp64", e)
    for k in FLUX:
        assert e[k] <= 0.1, k
    for k in HR:
        assert e[k] <= 0.01, k


@pytest.mark.parametrize("nlay,icld,seed", LIMITS)
def test_host_abi_layer_limits(gpu_lib, nlay, icld, seed):
    cols = _cols(nlay, icld, seed)
    print(f"host nlay={nlay} icld={icld} seed={seed}", flush=True)
    mc, _ = ob.run_mcica(cols)
    got_mc = gpu_lib.mcica_subcol_sw(**mc_args(cols))
    for k in mc:
        assert np.array_equal(got_mc[k], mc[k]), k
    got = gpu_lib.rrtmg_sw(**sw_args(cols, mc))
    ref = ob.run_rrtmg(cols, mc, icld=icld)
    print(_err(got, ref))
    assert_close(got, ref, f"nlay={nlay}")


def _adjoint(api, cols, d, st, sh, ckpt):
    import torch
    g = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ckpt)
    torch.cuda.synchronize()
    return g


def _toa_and_hr(api, cols):
    import torch
    out = api.mcica_rrtmg_sw(cols, api.device_inputs(cols), precision="exact")
    torch.cuda.synchronize()
    return api.to_host(out["swuflx"])[:, cols.nlay], api.to_host(out["swhr"])


@pytest.mark.parametrize("nlay,seed,step", [(4, 14, 1e-4), (256, 48, 1e-6)])
def test_adjoint_layer_limits(gpu_lib, nlay, seed, step):
    """Checkpointing recomputes the stored sweep states exactly, and a repeated call returns the
    same bits (the accumulations into shared memory add in a fixed order). The gradients of
    J1 = TOA upward flux and J2 = sum_k w_k swhr_k agree within 1e-4 relative, in every column,
    with finite differences of the exact-exponential forward along a random relative direction
    (rrtmg_sw_amd.gradcheck: the central difference, or where the step crosses a kink of the
    piecewise-linear tables a one-sided Richardson quotient, at the step or at step / 8, each
    with its rounding allowance). The weights w_k = u_k pdp_k / max pdp (u uniform in [0, 1])
    keep J2 well conditioned at any layer count: the heating rate of a layer is its flux
    divergence times heatfac / pdp, and in the thin top layers of a 256-layer column
    (pdp ~ 5e-4 hPa) that factor lifts the fluxes' rounding above a difference of the plain
    sum. At 256 layers the step is 1e-6 |x|, so that few steps cross a 15 K temperature cell
    boundary."""
    import torch
    from rrtmg_sw_amd import gradcheck
    api = gpu_lib
    cols = make_columns(16, nlay, icld=5, iaer=10, seed=seed)
    N, L = cols.ncol, cols.nlay
    print(f"adjoint nlay={nlay} seed={seed}", flush=True)
    d = api.device_inputs(cols)
    st = torch.linspace(0.5, 1.5, N, dtype=torch.float64, device="cuda")
    sh = torch.rand(L, N, dtype=torch.float64, device="cuda",
                    generator=torch.Generator("cuda").manual_seed(9))
    ref = _adjoint(api, cols, d, st, sh, 0)
    for k in [0] + sorted({1, 7, 16, L} & set(range(1, L + 1))):   # 0: run to run
        g = _adjoint(api, cols, d, st, sh, k)
        for name in ref:
            assert torch.equal(ref[name], g[name]), (k, name)
    rng = np.random.default_rng(seed)
    pdp = cols.plev[:, :-1] - cols.plev[:, 1:]
    w = rng.uniform(0.0, 1.0, (N, L)) * pdp / pdp.max()
    dirs = {k: step * getattr(cols, k) * rng.uniform(-1, 1, getattr(cols, k).shape)
            for k in KINDS + ALB}
    objs = {1: (np.ones(N), np.zeros((N, L))), 2: (np.zeros(N), w)}

    def adjoint(s1, s2):
        g = _adjoint(api, cols, d, api.to_device(s1), api.to_device(s2), 0)
        return {k: api.to_host(v) for k, v in g.items()}

    for obj, (rel, n1, n8) in gradcheck.directional(api, cols, dirs, objs, adjoint).items():
        print(f"J{obj} nlay={nlay}: sorted relative errors " + " ".join(f"{x:.1e}" for x in np.sort(rel))
              + f"; {n1} one-sided, {n8} at step / 8")
        assert rel.max() <= 1e-4, (obj, rel)


def test_single_column(gpu_lib):
    """One column: the fused forward and the non-McICA driver against the oracle, and the
    adjoint equal to column 0 of a 24-column call (columns are independent and the KISS streams
    depend on the column's pressures only)."""
    import torch
    api = gpu_lib
    many = make_columns(24, 60, icld=2, iaer=10, seed=21)
    cols = many.subset(slice(0, 1))
    print("single column", flush=True)
    got = _fused(api, cols)
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc)
    print(_err(got, ref))
    assert_close(got, ref, "ncol=1")
    full = _fused(api, many)
    for k in FLUX + HR:
        assert np.array_equal(full[k][:1], got[k]), k
    oc = overcast(cols)
    got = api.rrtmg_sw_nomcica(**nm_args(oc))
    ref, _, _ = ob.run_nomcica(oc)
    assert_close(got, ref, "nomcica ncol=1")
    seeds = lambda c: (torch.ones(c.ncol, dtype=torch.float64, device="cuda"),   # noqa: E731
                       torch.ones(c.nlay, c.ncol, dtype=torch.float64, device="cuda"))
    g1 = _adjoint(api, cols, api.device_inputs(cols), *seeds(cols), 0)
    gn = _adjoint(api, many, api.device_inputs(many), *seeds(many), 0)
    for k in g1:   # torch shapes (nlay, ncol) and (4, ncol)
        assert torch.equal(g1[k], gn[k][:, :1]), k


def test_no_columns(gpu_lib):
    """ncol = 0: every entry returns at once with no error (the reference's column loops run
    zero times); the written-back flags keep the reference's clamps."""
    import torch
    api = gpu_lib
    cols = make_columns(4, 30, icld=2, iaer=10, seed=2).subset(slice(0, 0))
    assert cols.ncol == 0
    mc = ob.empty_mcica(cols)
    out = api.rrtmg_sw(**{**sw_args(cols, mc), "icld": 7, "iaer": 3})
    assert (out["icld"], out["iaer"]) == (2, 0)
    assert out["swuflx"].shape == (0, 31)
    out = api.rrtmg_sw_nomcica(**nm_args(cols))
    assert out["swhr"].shape == (0, 30)
    out = api.mcica_subcol_sw(**mc_args(cols))
    assert out["cldfmcl"].size == 0
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d)
    assert out["swuflx"].numel() == 0
    g = api.mcica_rrtmg_sw_adjoint(cols, d, torch.zeros(0, dtype=torch.float64, device="cuda"),
                                   torch.zeros(30, 0, dtype=torch.float64, device="cuda"))
    torch.cuda.synchronize()
    assert g["tlay"].numel() == 0


@pytest.mark.parametrize("nlay", [3, 257])
def test_layer_count_rejected(gpu_lib, nlay):
    """Outside 4..256 layers every column entry returns RRTMGSW_ERR_ARGUMENT before any work."""
    import torch
    api = gpu_lib
    cols = make_columns(2, nlay, icld=2, iaer=0, seed=3)
    mc = ob.empty_mcica(cols)
    d = api.device_inputs(cols)
    st = torch.ones(2, dtype=torch.float64, device="cuda")
    sh = torch.ones(nlay, 2, dtype=torch.float64, device="cuda")
    calls = {
        "rrtmg_sw": lambda: api.rrtmg_sw(**sw_args(cols, mc)),
        "nomcica": lambda: api.rrtmg_sw_nomcica(**nm_args(cols)),
        "fused": lambda: api.mcica_rrtmg_sw(cols, d),
        "adjoint": lambda: api.mcica_rrtmg_sw_adjoint(cols, d, st, sh),
    }
    # mcica_subcol_sw takes fewer than 4 layers with the Mersenne Twister, not with KISS
    calls["mcica_subcol_sw (KISS)"] = lambda: api.mcica_subcol_sw(**mc_args(cols))
    for name, fn in calls.items():
        with pytest.raises(api.RRTMGSWError) as e:
            fn()
        assert e.value.code == 2, name


def test_reference_stops(gpu_lib):
    """The reference's STOP conditions come back as error codes with its messages."""
    import torch
    api = gpu_lib
    cols = make_columns(6, 40, icld=2, iaer=0, seed=31)
    with pytest.raises(api.RRTMGSWError, match="MCICA_SUBCOL: INVALID ICLD") as e:
        api.mcica_subcol_sw(**{**mc_args(cols), "icld": 6})
    assert e.value.code == 10
    bad = cols.subset(np.arange(6))
    bad.icld = 7
    with pytest.raises(api.RRTMGSWError, match="INVALID ICLD") as e:
        api.mcica_rrtmg_sw(bad, api.device_inputs(bad))
    assert e.value.code == 10
    bad = cols.subset(np.arange(6))
    bad.inflgsw = 1
    with pytest.raises(api.RRTMGSWError, match="INFLAG = 1 OPTION NOT AVAILABLE WITH MCICA") as e:
        api.mcica_rrtmg_sw(bad, api.device_inputs(bad))
    assert e.value.code == 3
    # the two lowest layers of column 4 out of order: the KISS seed test fails on the GPU
    bad = cols.subset(np.arange(6))
    bad.play[4, 0], bad.play[4, 1] = bad.play[4, 1], bad.play[4, 0]
    with pytest.raises(api.RRTMGSWError, match="REQUIRES PMID FROM BOTTOM FOUR LAYERS") as e:
        api.mcica_subcol_sw(**mc_args(bad))
    assert e.value.code == 11
    with pytest.raises(api.RRTMGSWError, match="REQUIRES PMID FROM BOTTOM FOUR LAYERS") as e:
        api.mcica_rrtmg_sw(bad, api.device_inputs(bad))
    assert e.value.code == 11
    # particle sizes outside the table bounds in a layer with cloudy sub-columns: the message
    # names the parameterization, as the oracle's (and the reference's) STOP does
    mc, _ = ob.run_mcica(cols)
    cloudy = mc["cldfmcl"].any(axis=0)
    ice = np.argwhere((cols.ciwp > 0) & cloudy)[0]
    liq = np.argwhere((cols.clwp > 0) & cloudy)[0]
    for what, iceflg, at, value, msg in (
            ("rei", 3, ice, 150.0, "ICE GENERALIZED EFFECTIVE SIZE OUT OF BOUNDS"),
            ("rei", 2, ice, 135.0, "ICE RADIUS OUT OF BOUNDS"),
            ("rei", 1, ice, 10.0, "ICE RADIUS OUT OF BOUNDS"),
            ("rel", 3, liq, 70.0, "LIQUID EFFECTIVE RADIUS OUT OF BOUNDS")):
        bad = cols.subset(np.arange(6))
        bad.iceflgsw = iceflg
        getattr(bad, what)[at[0], at[1]] = value
        with pytest.raises(RuntimeError, match=msg):
            ob.run_rrtmg(bad, ob.run_mcica(bad)[0])
        with pytest.raises(api.RRTMGSWError, match=msg) as e:
            api.mcica_rrtmg_sw(bad, api.device_inputs(bad))
        assert e.value.code == 4 and e.value.message == msg, (what, iceflg)
        # the non-McICA driver's cldprop_sw tests the same bounds on its overcast layers
        oc = overcast(bad)
        with pytest.raises(api.RRTMGSWError, match=msg) as e:
            api.rrtmg_sw_nomcica(**nm_args(oc))
        assert e.value.code == 4
    torch.cuda.synchronize()
    # the library stays usable after each of these
    got = _fused(api, cols)
    mc, _ = ob.run_mcica(cols)
    assert_close(got, ob.run_rrtmg(cols, mc), "after errors")


def test_inflag_one_without_cloudy_cells(gpu_lib):
    """inflag = 1 stops only in a cell that passes the cloud test (src/rrtmg_sw_cldprmc.f90:145-166):
    cloud-free columns (cloud fraction 0, water paths left as they are) run through the fused entry,
    mcica_subcol_sw + rrtmg_sw and the adjoint and agree with the oracle; one cloudy layer stops."""
    import torch
    api = gpu_lib
    cols = make_columns(6, 40, icld=2, iaer=0, seed=32)
    cols.cldfrac = np.zeros_like(cols.cldfrac, order="F")
    cols.inflgsw = 1
    mc, _ = ob.run_mcica(cols)
    assert not mc["cldfmcl"].any()
    ref = ob.run_rrtmg(cols, mc)
    out = api.mcica_rrtmg_sw(cols, api.device_inputs(cols))
    torch.cuda.synchronize()
    assert_close({k: api.to_host(v) for k, v in out.items()}, ref, "fused, inflag 1")
    gmc = api.mcica_subcol_sw(**mc_args(cols))
    assert_close(api.rrtmg_sw(**sw_args(cols, gmc)), ref, "host ABI, inflag 1")
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
    g = api.mcica_rrtmg_sw_adjoint(cols, api.device_inputs(cols), st, sh)
    torch.cuda.synchronize()
    assert all(torch.isfinite(v).all() for v in g.values())
    bad = cols.subset(np.arange(6))
    f = bad.cldfrac.copy()
    f[3, np.argmax(bad.ciwp[3] + bad.clwp[3])] = 1.0   # a layer holding condensate
    bad.cldfrac = np.asfortranarray(f)
    with pytest.raises(RuntimeError, match="INFLAG = 1 OPTION NOT AVAILABLE WITH MCICA"):
        ob.run_rrtmg(bad, ob.run_mcica(bad)[0])
    with pytest.raises(api.RRTMGSWError, match="INFLAG = 1 OPTION NOT AVAILABLE WITH MCICA") as e:
        api.mcica_rrtmg_sw(bad, api.device_inputs(bad))
    assert e.value.code == 3


def test_finalize_and_reinit(gpu_lib):
    """rrtmgsw_finalize frees the tables and buffers; calls then return RRTMGSW_ERR_NOT_INITIALIZED
    until rrtmgsw_init, after which results are bit-for-bit those before."""
    api = gpu_lib
    cols = make_columns(16, 30, icld=2, iaer=10, seed=81)
    before = _fused(api, cols)
    from rrtmg_sw_amd import _lib
    try:
        assert _lib.lib().rrtmgsw_finalize() == 0
        with pytest.raises(api.RRTMGSWError) as e:
            api.mcica_rrtmg_sw(cols, api.device_inputs(cols))
        assert e.value.code == 1
    finally:
        api.rrtmg_sw_ini(1004.64)
    after = _fused(api, cols)
    for k in before:
        assert np.array_equal(before[k], after[k]), k


def test_night_and_grazing_sun(gpu_lib):
    """coszen at and below zero and down to 1e-12 (the driver clamps it to 1e-10,
    src/rrtmg_sw_rad.f90:665-666): the fused and host-array McICA entries and the non-McICA driver
    against the oracle, and the adjoint finite everywhere (exp(-tau/mu0) underflows to zero)."""
    import torch
    api = gpu_lib
    cols = make_columns(8, 60, icld=2, iaer=10, seed=91)
    cols.coszen = np.asfortranarray(np.array([-0.5, 0.0, 1e-12, 1e-9, 1e-6, 1e-3, 0.02, 1.0]))
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc)
    got = _fused(api, cols)
    print("fused", _err(got, ref))
    assert_close(got, ref, "night fused")
    assert_close(api.rrtmg_sw(**sw_args(cols, mc)), ref, "night host")
    oc = overcast(cols)
    nref, _, _ = ob.run_nomcica(oc)
    assert_close(api.rrtmg_sw_nomcica(**nm_args(oc)), nref, "night nomcica")
    d = api.device_inputs(cols)
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
    for ck in (0, 16):
        g = _adjoint(api, cols, d, st, sh, ck)
        for k, v in g.items():
            assert torch.isfinite(v).all(), (ck, k)


def test_extreme_atmospheres(gpu_lib):
    """Temperatures off the reference grid (140-360 K), water vapour 0 to 0.06, ozone 0 to 1e-4,
    a 0.001 hPa top (heating rates up to 2000 K/day in the thinnest layers): the fused and
    host-array McICA entries and the non-McICA driver against the oracle, the adjoint finite."""
    import torch
    from rrtmg_sw_amd.workloads import extreme
    api = gpu_lib
    cols = extreme()
    mc, _ = ob.run_mcica(cols)
    ref = ob.run_rrtmg(cols, mc)
    got = _fused(api, cols)
    print("fused", _err(got, ref))
    assert_close(got, ref, "extreme fused")
    assert_close(api.rrtmg_sw(**sw_args(cols, mc)), ref, "extreme host")
    oc = overcast(cols)
    nref, _, _ = ob.run_nomcica(oc)
    assert_close(api.rrtmg_sw_nomcica(**nm_args(oc)), nref, "extreme nomcica")
    d = api.device_inputs(cols)
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
    for ck in (0, 16):
        g = _adjoint(api, cols, d, st, sh, ck)
        for k, v in g.items():
            assert torch.isfinite(v).all(), (ck, k)


@pytest.mark.parametrize("case", ["night", "extreme"])
def test_precision_forms_on_edges(gpu_lib, case):
    """The mixed FP32/FP64 form within 0.1 W/m2 and 0.01 K/day of the FP64 form, and the
    exact-exponential form finite, on the night / grazing-sun and the extreme columns."""
    from rrtmg_sw_amd.workloads import extreme
    api = gpu_lib
    if case == "night":
        cols = make_columns(8, 60, icld=2, iaer=10, seed=91)
        cols.coszen = np.asfortranarray(np.array([-0.5, 0.0, 1e-12, 1e-9, 1e-6, 1e-3, 0.02, 1.0]))
    else:
        cols = extreme()
    f64 = _fused(api, cols)
    mixed = _fused(api, cols, "mixed")
    e = _err(mixed, f64)
    print(case, "mixed vs fp64", e)
    for k in FLUX:
        assert e[k] <= 0.1, k
    for k in HR:
        assert e[k] <= 0.01, k
    exact = _fused(api, cols, "exact")
    for k, v in exact.items():
        assert np.isfinite(v).all(), k


@pytest.mark.parametrize("icld,irng", [(1, 1), (3, 1), (4, 0)])
def test_adjoint_other_overlaps(gpu_lib, icld, irng):
    """The adjoint with random (1), maximum (3) and exponential (4) overlap, and with the Mersenne
    Twister generator: store and checkpoint-7 bit-identical, the TOA-flux directional derivative
    within the bar of the other adjoint tests."""
    import torch
    api = gpu_lib
    cols = make_columns(16, 40, icld=icld, iaer=10, seed=60 + icld)
    cols.irng = irng
    N, L = cols.ncol, cols.nlay
    print(f"adjoint icld={icld} irng={irng}", flush=True)
    d = api.device_inputs(cols)
    one = torch.ones(N, dtype=torch.float64, device="cuda")
    zero_hr = torch.zeros(L, N, dtype=torch.float64, device="cuda")
    a = _adjoint(api, cols, d, one, zero_hr, 0)
    b = _adjoint(api, cols, d, one, zero_hr, 7)
    for k in a:
        assert torch.equal(a[k], b[k]), k
    rng = np.random.default_rng(icld)
    dirs = {k: 1e-4 * getattr(cols, k) * rng.uniform(-1, 1, getattr(cols, k).shape)
            for k in KINDS + ALB}
    plus, minus = cols.subset(np.arange(N)), cols.subset(np.arange(N))
    for k, dx in dirs.items():
        getattr(plus, k)[...] += dx
        getattr(minus, k)[...] -= dx
    up_p, _ = _toa_and_hr(api, plus)
    up_m, _ = _toa_and_hr(api, minus)
    g = {k: api.to_host(v) for k, v in a.items()}
    lin = sum((g[k] * dirs[k]).sum(axis=1) for k in KINDS)
    lin = lin + sum(g["alb"][:, m] * dirs[k] for m, k in enumerate(ALB))
    rel = np.abs(lin - (up_p - up_m) / 2) / np.abs((up_p - up_m) / 2)
    print("J1: sorted relative errors " + " ".join(f"{x:.1e}" for x in np.sort(rel)))
    assert np.quantile(rel, 0.9) <= 1e-4 and rel.max() <= 1e-3, rel
