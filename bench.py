#!/usr/bin/env python3
"""Throughput of the GCM-mode McICA shortwave path on one MI355X.

Workload (one step): mcica_subcol_sw + rrtmg_sw on 8192 columns x 72 layers, McICA maximum-random
overlap (icld = 2) with the KISS generator, synthetic inputs of rrtmg_sw_amd.workloads
"mcica8192x72", all inputs resident in HBM before the timed region. The timed call is the fused
device entry point (rrtmgsw_mcica_sw_dev, FP64). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed] [--no-cpu]
    python bench.py --adjoint [--workload expran65536x137]

Both precision forms are timed in the same run; `value` is the one --precision names (FP64 by
default), and `forms` carries both, their speed ratio and the mixed form's largest error
against the FP64 form (fluxes W/m2, heating rates K/day). `adjoint` carries the adjoint's cost
against the forward on the 4096 x 60 exponential-random-overlap + aerosol workload; `nomcica`
the non-McICA driver on the same columns made overcast; `mt` the fused call with the Mersenne Twister
generator; `host_abi` the reference's own two
calls on host arrays (mcica_subcol_sw, then rrtmg_sw on the materialised sub-column arrays),
PCIe copies included; `cpu_baseline_all_cores` the CPU leg on 16 worker processes over the
whole workload.

--adjoint prints instead one line for the adjoint on a large workload (default 65536 x 137):
columns/s of the forward and of forward + adjoint, with full storage of every layer's terms
and with a checkpoint every 16 layers, and each mode's work-area footprint in HBM.
"""
import argparse
import ctypes as C
import gc
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the timed kernel on mcica8192x72, from rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE over this script, each scaled by what the counter reads for a 2 GiB device copy
# (x2.0 and x1.0); profiles/pmc_rt_kernel_adj_kernel_mcica8192x72_expran4096x60.md.
MEASURED_TRAFFIC = {"fp64": {"read": 2.064e9, "write": 1.911e9},
                    "mixed": {"read": 1.623e9, "write": 1.521e9}}


def cpu_baseline(cols_full, budget_s=12.0, chunk=16):
    """Time the compiled reference (oracle/_ref) -- mcica_subcol_sw + rrtmg_sw -- on consecutive
    column chunks of the same workload on one host core, until `budget_s` seconds of work. The
    reference keeps its (112, ncol, nlay) sub-column work arrays on the stack, so it is called
    on 16-column chunks, as a GCM calls it per chunk."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    kind = "reference" if ob.ref_available() else "port"
    which = "ref" if kind == "reference" else "orc"
    if kind == "reference":
        ob.ref()
    else:
        ob.orc()
    done, t = 0, 0.0
    while t < budget_s and done < cols_full.ncol:
        cols = cols_full.subset(slice(done, min(done + chunk, cols_full.ncol)))
        t0 = time.perf_counter()
        mc, _ = ob.run_mcica(cols, which)
        ob.run_rrtmg(cols, mc, which=which)
        t += time.perf_counter() - t0
        done += cols.ncol
    return {"value": done / t, "unit": "columns/s", "cores": 1, "kind": kind,
            "sample": f"{done} columns of the 8192x72 McICA workload (mcica_subcol_sw + rrtmg_sw, "
                      f"chunks of {chunk}), {t:.1f} s on one host core"}


_CPU = {}


def _cpu_worker_init(name, barrier):
    _CPU["barrier"] = barrier
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from rrtmg_sw_amd.workloads import workload
    _CPU["ob"] = ob
    _CPU["which"] = "ref" if ob.ref_available() else "orc"
    ob.ref() if _CPU["which"] == "ref" else ob.orc()
    _CPU["cols"] = workload(name)


def _cpu_worker_ready(_):
    _CPU["barrier"].wait(timeout=300)   # one task per worker: all of them initialised
    return 0


def _cpu_worker_block(lo_hi, chunk=16):
    ob, which, full = _CPU["ob"], _CPU["which"], _CPU["cols"]
    lo, hi = lo_hi
    for c0 in range(lo, hi, chunk):
        cols = full.subset(slice(c0, min(c0 + chunk, hi)))
        mc, _ = ob.run_mcica(cols, which)
        ob.run_rrtmg(cols, mc, which=which)
    return hi - lo


def cpu_baseline_all_cores(name="mcica8192x72", procs=16, reps=3):
    """The same CPU leg on `procs` worker processes (the reference has no threading), each on a
    contiguous block of the workload's columns in 16-column chunks; KISS streams are per column,
    so the blocks compute exactly the columns of a single run. Workers are forked before this
    process touches the GPU; the time is the wall clock of the whole workload once every worker
    has loaded the library and built the inputs."""
    import multiprocessing as mp
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    from rrtmg_sw_amd.workloads import WORKLOADS
    n = WORKLOADS[name]["ncol"]
    procs = max(1, min(procs, os.cpu_count() or 1))
    edges = [n * i // procs for i in range(procs + 1)]
    ctx = mp.get_context("fork")
    barrier = ctx.Barrier(procs)
    with ctx.Pool(procs, initializer=_cpu_worker_init, initargs=(name, barrier)) as pool:
        pool.map(_cpu_worker_ready, range(procs), chunksize=1)
        walls = []
        for _ in range(reps):
            t0 = time.perf_counter()
            done = sum(pool.map(_cpu_worker_block, list(zip(edges[:-1], edges[1:])), chunksize=1))
            walls.append(time.perf_counter() - t0)
    t = sorted(walls)[len(walls) // 2]
    return {"value": done / t, "unit": "columns/s", "cores": procs,
            "kind": "reference" if ob.ref_available() else "port",
            "sample": f"all {done} columns of {name} (mcica_subcol_sw + rrtmg_sw, chunks of 16) "
                      f"on {procs} worker processes, one contiguous block each; median of {reps} "
                      f"runs, {t:.2f} s wall"}


def _time(fn, steps, warmup):
    """Per-call milliseconds of `fn` over `steps` calls after `warmup`, synchronised."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps * 1e3


def adjoint_ratio(args, name="expran4096x60"):
    """Adjoint (total-sky TOA upward flux + heating-rate seeds, all five inputs + albedos)
    against the forward on the same columns."""
    import torch
    from rrtmg_sw_amd import api
    from rrtmg_sw_amd.workloads import workload
    cols = workload(name)
    d = api.device_inputs(cols)
    st = torch.ones(cols.ncol, dtype=torch.float64, device="cuda")
    sh = torch.ones(cols.nlay, cols.ncol, dtype=torch.float64, device="cuda")
    res = {"workload": f"{name}: icld=5 KISS, iaer=10, {cols.ncol} columns x {cols.nlay} layers"}
    for p in ("fp64", "exact"):
        out = api.mcica_rrtmg_sw(cols, d, precision=p)
        res[f"forward_{p}_ms"] = _time(lambda: api.mcica_rrtmg_sw(cols, d, precision=p, out=out),
                                       args.steps, args.warmup)
    for ck in (0, 16):
        g = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ck)
        res[f"adjoint_ckpt{ck}_ms"] = _time(
            lambda: api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, out=g, ckpt=ck), args.steps,
            args.warmup)
    res["adjoint_over_forward_fp64"] = res["adjoint_ckpt0_ms"] / res["forward_fp64_ms"]
    res["adjoint_over_forward_exact"] = res["adjoint_ckpt0_ms"] / res["forward_exact_ms"]
    return res


def nomcica_form(args, name="overcast8192x72"):
    """The non-McICA driver (rrtmgsw_sw_nomcica_dev: layer cloud fractions, cldprop_sw optics) on
    the same columns with every cloudy layer overcast, FP64, device-resident inputs."""
    from rrtmg_sw_amd import api
    from rrtmg_sw_amd.workloads import workload
    cols = workload(name)
    d = api.device_inputs(cols)
    kw = dict(ncol=cols.ncol, nlay=cols.nlay, icld=cols.icld, iaer=cols.iaer, play=d["play"],
              plev=d["plev"], tlay=d["tlay"], tlev=None, tsfc=None, h2ovmr=d["h2ovmr"],
              o3vmr=d["o3vmr"], co2vmr=d["co2vmr"], ch4vmr=d["ch4vmr"], n2ovmr=None,
              o2vmr=d["o2vmr"], asdir=d["asdir"], asdif=d["asdif"], aldir=d["aldir"],
              aldif=d["aldif"], coszen=d["coszen"], adjes=cols.adjes, dyofyr=cols.dyofyr,
              scon=cols.scon, isolvar=cols.isolvar, inflgsw=cols.inflgsw,
              iceflgsw=cols.iceflgsw, liqflgsw=cols.liqflgsw, cldfr=d["cldfrac"],
              taucld=d["tauc"], ssacld=d["ssac"], asmcld=d["asmc"], fsfcld=d["fsfc"],
              cicewp=d["ciwp"], cliqwp=d["clwp"], reice=d["rei"], reliq=d["rel"], tauaer=None,
              ssaaer=None, asmaer=None, ecaer=None)
    ms = _time(lambda: api.rrtmg_sw_nomcica(**kw), args.steps, args.warmup)
    # per call: six output tensors allocated and zeroed, the kernel, the status read-back
    return {"workload": f"{name}: rrtmg_sw of the non-McICA driver (cldprop_sw, spcvrt_sw), "
                        f"{cols.ncol} columns x {cols.nlay} layers, every cloudy layer overcast",
            "ms_per_step": ms, "columns_per_s": cols.ncol / (ms * 1e-3),
            "kernel": "rt_kernel<0, 2>"}


def mt_form(cols, d, steps=3, warmup=1):
    """The same fused call with the Mersenne Twister generator (irng = 1): one global MT19937
    stream per call, generated in order by one workgroup, then the same kernel."""
    from rrtmg_sw_amd import api
    irng = cols.irng
    cols.irng = 1
    try:
        out = api.mcica_rrtmg_sw(cols, d)
        ms = _time(lambda: api.mcica_rrtmg_sw(cols, d, out=out), steps, warmup)
    finally:
        cols.irng = irng
    return {"workload": f"mcica8192x72 with irng = 1 (one MT19937 stream per call)",
            "ms_per_step": ms, "columns_per_s": cols.ncol / (ms * 1e-3)}


def host_abi_form(cols, steps=3, warmup=1):
    """The reference's two calls on host arrays, as a GCM makes them through the drop-in:
    rrtmgsw_mcica_subcol (profiles copied in; the seven (112, ncol, nlay) sub-column arrays and
    two (ncol, nlay) arrays copied back) then rrtmgsw_sw (those arrays and the profiles copied
    in, fluxes and heating rates copied back), each synchronous, on pageable host arrays
    allocated once, as a caller's arrays are. Wall time per call; PCIe-inclusive, so never the
    bench value."""
    from rrtmg_sw_amd import api
    N, L, G = cols.ncol, cols.nlay, api.NGPT
    mc_kw = dict(iplon=1, ncol=N, nlay=L, icld=cols.icld, permuteseed=cols.permuteseed,
                 irng=cols.irng, play=cols.play, cldfrac=cols.cldfrac, ciwp=cols.ciwp,
                 clwp=cols.clwp, rei=cols.rei, rel=cols.rel, tauc=cols.tauc, ssac=cols.ssac,
                 asmc=cols.asmc, fsfc=cols.fsfc, alpha=cols.alpha)
    mc = api.mcica_subcol_sw(**mc_kw)

    def sw_kw(mc):
        return dict(ncol=N, nlay=L, icld=cols.icld, iaer=cols.iaer, play=cols.play,
                    plev=cols.plev, tlay=cols.tlay, tlev=cols.tlev, tsfc=cols.tsfc,
                    h2ovmr=cols.h2ovmr, o3vmr=cols.o3vmr, co2vmr=cols.co2vmr,
                    ch4vmr=cols.ch4vmr, n2ovmr=cols.n2ovmr, o2vmr=cols.o2vmr, asdir=cols.asdir,
                    asdif=cols.asdif, aldir=cols.aldir, aldif=cols.aldif, coszen=cols.coszen,
                    adjes=cols.adjes, dyofyr=cols.dyofyr, scon=cols.scon, isolvar=cols.isolvar,
                    inflgsw=cols.inflgsw, iceflgsw=cols.iceflgsw, liqflgsw=cols.liqflgsw,
                    cldfmcl=mc["cldfmcl"], taucmcl=mc["taucmcl"], ssacmcl=mc["ssacmcl"],
                    asmcmcl=mc["asmcmcl"], fsfcmcl=mc["fsfcmcl"], ciwpmcl=mc["ciwpmcl"],
                    clwpmcl=mc["clwpmcl"], reicmcl=mc["reicmcl"], relqmcl=mc["relqmcl"],
                    tauaer=cols.tauaer, ssaaer=cols.ssaaer, asmaer=cols.asmaer, ecaer=cols.ecaer)

    out = api.rrtmg_sw(**sw_kw(mc))
    t_sub = t_sw = 0.0
    for i in range(warmup + steps):
        t0 = time.perf_counter()
        api.mcica_subcol_sw(**mc_kw, out=mc)
        t1 = time.perf_counter()
        api.rrtmg_sw(**sw_kw(mc), out=out)
        t2 = time.perf_counter()
        if i >= warmup:
            t_sub += t1 - t0
            t_sw += t2 - t1
    t_sub, t_sw = t_sub / steps * 1e3, t_sw / steps * 1e3
    big = 8.0 * (7 * G * N * L + 2 * N * L)
    return {"workload": f"mcica8192x72 on host arrays: rrtmgsw_mcica_subcol then rrtmgsw_sw, "
                        f"{steps} timed pairs after {warmup}",
            "ms_mcica_subcol": t_sub, "ms_rrtmg_sw": t_sw, "ms_per_step": t_sub + t_sw,
            "columns_per_s": N / ((t_sub + t_sw) * 1e-3),
            "subcolumn_bytes_each_way": big,
            "note": "PCIe-inclusive: the sub-column arrays go device -> host and back; the fused "
                    "device entry avoids both copies"}


def bench_adjoint(args):
    """Forward and forward + adjoint on a large workload, store vs checkpoint-16."""
    import torch
    from rrtmg_sw_amd import api, _lib
    from rrtmg_sw_amd.workloads import workload
    name = args.workload if args.workload != "mcica8192x72" else "expran65536x137"
    cols = workload(name)
    d = api.device_inputs(cols)
    N, L = cols.ncol, cols.nlay
    st = torch.ones(N, dtype=torch.float64, device="cuda")
    sh = torch.ones(L, N, dtype=torch.float64, device="cuda")
    steps, warmup = max(2, args.steps // 4), 1
    out = api.mcica_rrtmg_sw(cols, d, precision="fp64")
    fwd = _time(lambda: api.mcica_rrtmg_sw(cols, d, precision="fp64", out=out), steps, warmup)
    line = {"metric": "adjoint columns/s (forward + adjoint), store vs checkpoint-16",
            "unit": "columns/s", "n_gpus": 1, "steps": steps, "warmup": warmup,
            "higher_is_better": True, "data": "synthetic", "dtype": "f64",
            "config": {"workload": f"{name}: icld=5 KISS, iaer=10, {N} columns x {L} layers"},
            "forward_fp64": {"ms": fwd, "columns_per_s": N / (fwd * 1e-3)}}
    for ck in (0, 16):
        g = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ck)
        adj = _time(lambda: api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, out=g, ckpt=ck), steps,
                    warmup)
        per_col = _lib.lib().rrtmgsw_adjoint_work_bytes(C.c_int(L), C.c_int(ck))
        chunk = min(N, max(64, (4 << 30) // per_col))   # columns per launch (library's rule)
        line[f"ckpt{ck}"] = {
            "adjoint_ms": adj, "forward_plus_adjoint_columns_per_s": N / ((fwd + adj) * 1e-3),
            "adjoint_over_forward": adj / fwd, "work_bytes_per_column": per_col,
            "work_gib_if_all_columns_at_once": per_col * N / 2 ** 30,
            "work_gib_allocated": per_col * chunk / 2 ** 30, "columns_per_launch": chunk}
        del g
        torch.cuda.empty_cache()
    line["value"] = line["ckpt16"]["forward_plus_adjoint_columns_per_s"]
    print(json.dumps(line), flush=True)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--workload", default="mcica8192x72")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--adjoint", action="store_true")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one GPU; the path shards "
              f"by columns with no collective (see DESIGN.md). --gpus {args.gpus} is not run.",
              file=sys.stderr)
        sys.exit(2)

    # the all-core CPU leg forks its workers before this process touches the GPU
    cpu_all = None
    if not args.no_cpu and not args.adjoint and args.workload == "mcica8192x72":
        cpu_all = cpu_baseline_all_cores(args.workload)

    import torch
    from rrtmg_sw_amd import api, flops
    from rrtmg_sw_amd.workloads import workload

    api.rrtmg_sw_ini(1004.64)
    gc.disable()   # a collector pause inside a timed region would be counted as GPU time
    if args.adjoint:
        return bench_adjoint(args)
    cols = workload(args.workload)
    d = api.device_inputs(cols)

    # cloudy (g, layer) cells of this workload, for the algorithmic FLOP count
    mc = api.mcica_subcol_sw(1, cols.ncol, cols.nlay, cols.icld, cols.permuteseed, cols.irng,
                             d["play"], d["cldfrac"], d["ciwp"], d["clwp"], d["rei"], d["rel"],
                             d["tauc"], d["ssac"], d["asmc"], d["fsfc"], d["alpha"])
    cloudy = float(mc["cldfmcl"].sum().item())
    del mc
    torch.cuda.synchronize()
    fl = flops.call_flops(cols.ncol, cols.nlay, cloudy)

    def timed(precision):
        """Warm-up, then args.steps calls bracketed by synchronisations; HIP events on the
        stream the kernels are launched on (torch's current stream)."""
        out = api.mcica_rrtmg_sw(cols, d, precision=precision)
        for _ in range(args.warmup):
            api.mcica_rrtmg_sw(cols, d, precision=precision, out=out)
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            api.mcica_rrtmg_sw(cols, d, precision=precision, out=out)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        return out, wall / args.steps * 1e3, e0.elapsed_time(e1) / args.steps

    other = "mixed" if args.precision == "fp64" else "fp64"
    out, ms, kern_ms = timed(args.precision)
    out_o, ms_o, kern_ms_o = timed(other)
    o64, omx = (out, out_o) if args.precision == "fp64" else (out_o, out)
    err = {k: float((omx[k] - o64[k]).abs().max().item()) for k in o64}
    achieved = fl / (kern_ms * 1e-3) / 1e12
    line = {
        "metric": "McICA columns/s at 72 layers, FP32 vs FP64 GPU at <=0.1 W/m2; adjoint/fwd ratio",
        "value": cols.ncol / (ms * 1e-3),
        "unit": "columns/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: mcica_subcol_sw + rrtmg_sw, icld=2 KISS, "
                               f"{cols.ncol} columns x {cols.nlay} layers, fused device entry",
                   "precision": args.precision},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": (sum(MEASURED_TRAFFIC[args.precision].values())
                                 if args.workload == "mcica8192x72" else None),
                     "traffic_read_write": MEASURED_TRAFFIC[args.precision]
                                           if args.workload == "mcica8192x72" else None,
                     "kernel": f"rt_kernel<{0 if args.precision == 'fp64' else 1}, 1>",
                     "note": "FP64 VALU work (no dense contraction): algorithmic FLOPs of "
                             "rrtmg_sw_amd/flops.py over the per-step event time on the launch "
                             "stream; MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s",
                     "flops_per_step": fl, "event_ms_per_step": kern_ms},
        "forms": {
            args.precision: {"columns_per_s": cols.ncol / (ms * 1e-3), "ms_per_step": ms,
                             "event_ms_per_step": kern_ms},
            other: {"columns_per_s": cols.ncol / (ms_o * 1e-3), "ms_per_step": ms_o,
                    "event_ms_per_step": kern_ms_o},
            "mixed_over_fp64_speed": (ms if args.precision == "fp64" else ms_o) /
                                     (ms_o if args.precision == "fp64" else ms),
            "mixed_max_abs_error_vs_fp64": err,
        },
    }
    line["adjoint"] = adjoint_ratio(args)
    line["nomcica"] = nomcica_form(args)
    if args.workload == "mcica8192x72":
        line["mt"] = mt_form(cols, d)
        del d
        torch.cuda.empty_cache()
        line["host_abi"] = host_abi_form(cols)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cols)
    if cpu_all is not None:
        line["cpu_baseline_all_cores"] = cpu_all
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
