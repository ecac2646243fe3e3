#!/usr/bin/env python3
"""Throughput of the GCM-mode McICA shortwave path on one MI355X.

Workload (one step): mcica_subcol_sw + rrtmg_sw on 8192 columns x 72 layers, McICA maximum-random
overlap (icld = 2) with the KISS generator, synthetic inputs of rrtmg_sw_amd.workloads
"mcica8192x72", all inputs resident in HBM before the timed region. The timed call is the fused
device entry point (rrtmgsw_mcica_sw_dev, FP64). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed] [--no-cpu] [--no-pmc]
    python bench.py --adjoint [--workload expran65536x137]

Both precision forms are timed in the same run; `value` is the one --precision names (FP64 by
default), and `forms` carries both, their speed ratio and each output's largest error against
the FP64 form and against the compiled reference run by the CPU leg (fluxes W/m2, heating rates
K/day). `roofline` is the timed kernel's: FLOPs counted by the oracle's operation-counting build
(rrtmg_sw_amd/data/flops_counted.json) over its event time, against the FP64 vector peak; its
`traffic` is measured in this run by two rocprofv3 --pmc passes (FETCH_SIZE, WRITE_SIZE) over a
child run of the timed call, each scaled by the same counter on a 1 GiB elementwise kernel
(--no-pmc, or no rocprofv3: the figures recorded under profiles/). `adjoint` carries the
adjoint's cost against the forward on the 4096 x 60 exponential-random-overlap + aerosol
workload; `adjoint_large` the same on 65536 x 137 (forward forms, store and checkpoint-16,
work-area footprint, directional finite-difference check on sampled columns); `nomcica` the
non-McICA driver on the same columns made overcast; `mt` the fused call with the Mersenne
Twister generator; `host_abi` the reference's own two calls on host arrays (mcica_subcol_sw,
then rrtmg_sw on the materialised sub-column arrays), PCIe copies included;
`cpu_baseline_all_cores` the CPU leg on 16 worker processes over the whole workload.

--adjoint prints instead one line for the adjoint on a large workload (default 65536 x 137).
"""
import argparse
import csv
import ctypes as C
import gc
import json
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 vector (and FP64 matrix) peak
PEAK_HBM_GBS = 8000.0

# HBM bytes per launch of the timed kernel on mcica8192x72 recorded with rocprofv3 --pmc
# FETCH_SIZE and WRITE_SIZE (the calibrated counters), used when this run cannot measure them;
# profiles/pmc_rt_kernel_mcica8192x72.md.
RECORDED_TRAFFIC = {"fp64": {"read": 2.064e9, "write": 1.911e9},
                    "mixed": {"read": 1.623e9, "write": 1.521e9}}
CALIB_DOUBLES = 1 << 27    # 1 GiB read and 1 GiB written by the calibration kernel


def kernel_only(args):
    """Child of measure_traffic: the timed call alone (one warm-up, two calls), then the
    calibration kernel y = 2 x over 1 GiB."""
    import torch
    from rrtmg_sw_amd import api
    from rrtmg_sw_amd.workloads import workload
    api.rrtmg_sw_ini(1004.64)
    cols = workload(args.workload)
    d = api.device_inputs(cols)
    out = api.mcica_rrtmg_sw(cols, d, precision=args.precision)
    for _ in range(2):
        api.mcica_rrtmg_sw(cols, d, precision=args.precision, out=out)
    torch.cuda.synchronize()
    x = torch.ones(CALIB_DOUBLES, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 2.0, out=y)
    torch.cuda.synchronize()
    print("kernel-only done", flush=True)


def measure_traffic(args):
    """HBM bytes per launch of the timed kernel (read, write), measured: one rocprofv3 --pmc pass
    per counter over a child run of this script (--kernel-only), the kernel's counter (KiB per
    dispatch, averaged over the dispatches after the first) scaled by 1 GiB over what the same
    counter reads for the calibration kernel. Runs before this process touches the GPU. Returns
    None when rocprofv3 is absent or a pass fails."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        return None
    got = {}
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        for ctr, key in (("FETCH_SIZE", "read"), ("WRITE_SIZE", "write")):
            cmd = [exe, "--pmc", ctr, "--kernel-include-regex", "rt_kernel|elementwise", "-f",
                   "csv", "-d", tmp, "-o", ctr, "--", sys.executable, os.path.abspath(__file__),
                   "--kernel-only", "--precision", args.precision, "--workload", args.workload]
            proc = subprocess.Popen(cmd, cwd="/tmp", env={**os.environ, "TMPDIR": "/tmp"},
                                    stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                                    start_new_session=True)
            try:
                rc = proc.wait(timeout=150)
            except subprocess.TimeoutExpired:
                os.killpg(proc.pid, signal.SIGKILL)
                proc.wait()
                return None
            path = os.path.join(tmp, f"{ctr}_counter_collection.csv")
            if rc != 0 or not os.path.exists(path):
                return None
            kern, calib = [], []
            with open(path) as f:
                for row in csv.DictReader(f):
                    if row["Counter_Name"] != ctr:
                        continue
                    v = float(row["Counter_Value"]) * 1024.0
                    (kern if "rt_kernel<" in row["Kernel_Name"] else calib).append(v)
            if len(kern) < 2 or not calib or calib[-1] <= 0:
                return None
            scale = 8.0 * CALIB_DOUBLES / calib[-1]
            got[key] = sum(kern[1:]) / len(kern[1:]) * scale
            got[key + "_scale"] = scale
    return got


def cpu_baseline(cols_full, budget_s=12.0, chunk=16):
    """Time the compiled reference (oracle/_ref) -- mcica_subcol_sw + rrtmg_sw -- on consecutive
    column chunks of the same workload on one host core, until `budget_s` seconds of work. The
    reference keeps its (112, ncol, nlay) sub-column work arrays on the stack, so it is called
    on 16-column chunks, as a GCM calls it per chunk. Returns the record and the reference's
    outputs of the columns it computed (the first `done`)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    kind = "reference" if ob.ref_available() else "port"
    which = "ref" if kind == "reference" else "orc"
    if kind == "reference":
        ob.ref()
    else:
        ob.orc()
    done, t, outs = 0, 0.0, []
    while t < budget_s and done < cols_full.ncol:
        cols = cols_full.subset(slice(done, min(done + chunk, cols_full.ncol)))
        t0 = time.perf_counter()
        mc, _ = ob.run_mcica(cols, which)
        o = ob.run_rrtmg(cols, mc, which=which)
        t += time.perf_counter() - t0
        outs.append(o)
        done += cols.ncol
    ref = {k: np.concatenate([o[k] for o in outs]) for k in outs[0] if k not in ("icld", "iaer")}
    return {"value": done / t, "unit": "columns/s", "cores": 1, "kind": kind,
            "sample": f"{done} columns of the 8192x72 McICA workload (mcica_subcol_sw + rrtmg_sw, "
                      f"chunks of {chunk}), {t:.1f} s on one host core"}, ref


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
    res["adjoint_ckpt16_over_forward_exact"] = res["adjoint_ckpt16_ms"] / res["forward_exact_ms"]
    from rrtmg_sw_amd import _lib
    for ck in (0, 16):
        res[f"work_bytes_per_column_ckpt{ck}"] = int(_lib.lib().rrtmgsw_adjoint_work_bytes(cols.nlay, ck))
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


def adjoint_large(args, name="expran65536x137", sample=16):
    """The large adjoint workload: the forward in its three forms, the adjoint with every
    layer's terms stored and with a checkpoint every 16 layers (time, ratio to the exact forward
    it differentiates and to the mixed forward, work-area footprint), and the directional
    derivative of each mode against finite differences of the exact forward on `sample`
    columns spread over the workload (rrtmg_sw_amd.gradcheck; objectives: the TOA upward flux
    and the heating-rate sum)."""
    import numpy as np
    import torch
    from rrtmg_sw_amd import _lib, api, gradcheck
    from rrtmg_sw_amd.workloads import workload
    cols = workload(name)
    d = api.device_inputs(cols)
    N, L = cols.ncol, cols.nlay
    st = torch.ones(N, dtype=torch.float64, device="cuda")
    sh = torch.ones(L, N, dtype=torch.float64, device="cuda")
    steps, warmup = max(2, args.steps // 4), 1
    res = {"workload": f"{name}: icld=5 KISS, iaer=10, {N} columns x {L} layers",
           "steps": steps, "warmup": warmup, "forward": {}}
    for p in ("fp64", "exact", "mixed"):
        out = api.mcica_rrtmg_sw(cols, d, precision=p)
        ms = _time(lambda: api.mcica_rrtmg_sw(cols, d, precision=p, out=out), steps, warmup)
        res["forward"][p] = {"ms": ms, "columns_per_s": N / (ms * 1e-3)}
        del out
    fx, fm = res["forward"]["exact"]["ms"], res["forward"]["mixed"]["ms"]
    for ck in (0, 16):
        g = api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, ckpt=ck)
        adj = _time(lambda: api.mcica_rrtmg_sw_adjoint(cols, d, st, sh, out=g, ckpt=ck), steps,
                    warmup)
        per_col = _lib.lib().rrtmgsw_adjoint_work_bytes(C.c_int(L), C.c_int(ck))
        chunk = min(N, max(64, (4 << 30) // per_col))   # columns per launch (library's rule)
        res[f"ckpt{ck}"] = {
            "adjoint_ms": adj, "adjoint_over_exact_forward": adj / fx,
            "adjoint_over_mixed_forward": adj / fm,
            "exact_forward_plus_adjoint_columns_per_s": N / ((fx + adj) * 1e-3),
            "work_bytes_per_column": per_col, "columns_per_launch": chunk,
            "work_gib_allocated": per_col * chunk / 2 ** 30,
            "work_gib_if_all_columns_at_once": per_col * N / 2 ** 30}
        del g
    del d, st, sh
    torch.cuda.empty_cache()
    sub = cols.subset(np.linspace(0, N - 1, sample).astype(int))
    rng = np.random.default_rng(13)
    dirs = {k: 1e-4 * getattr(sub, k) * rng.uniform(-1, 1, getattr(sub, k).shape)
            for k in gradcheck.KINDS + gradcheck.ALB}
    dsub = api.device_inputs(sub)
    objs = {"toa_up": (np.ones(sample), np.zeros((sample, L))),
            "hr_sum": (np.zeros(sample), np.ones((sample, L)))}
    fd = {}
    for ck in (0, 16):
        def adjoint(s_toa, s_hr):
            g = api.mcica_rrtmg_sw_adjoint(sub, dsub, api.to_device(s_toa), api.to_device(s_hr),
                                           ckpt=ck)
            torch.cuda.synchronize()
            return {k: api.to_host(v) for k, v in g.items()}
        for obj, (rel, n1, n8) in gradcheck.directional(api, sub, dirs, objs, adjoint).items():
            fd[f"ckpt{ck}_{obj}"] = {"median_rel": float(np.median(rel)), "max_rel": float(rel.max()),
                                     "one_sided": n1, "at_step_over_8": n8}
    res["fd_check"] = {"columns": sample, "inputs": list(gradcheck.KINDS + gradcheck.ALB),
                       "step": "1e-4 relative, random signs", "errors": fd}
    return res


def bench_adjoint(args):
    """One line for the large adjoint workload alone."""
    name = args.workload if args.workload != "mcica8192x72" else "expran65536x137"
    res = adjoint_large(args, name)
    line = {"metric": "adjoint columns/s (exact forward + adjoint), store vs checkpoint-16",
            "value": res["ckpt16"]["exact_forward_plus_adjoint_columns_per_s"],
            "unit": "columns/s", "n_gpus": 1, "steps": res["steps"], "warmup": res["warmup"],
            "higher_is_better": True, "data": "synthetic", "dtype": "f64",
            "config": {"workload": res["workload"]}, "adjoint_large": res}
    print(json.dumps(line), flush=True)


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--workload", default="mcica8192x72")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--no-pmc", action="store_true")
    p.add_argument("--adjoint", action="store_true")
    p.add_argument("--kernel-only", action="store_true", help=argparse.SUPPRESS)
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one GPU; the path shards "
              f"by columns with no collective (see DESIGN.md). --gpus {args.gpus} is not run.",
              file=sys.stderr)
        sys.exit(2)
    if args.kernel_only:
        return kernel_only(args)

    # the counter passes and the all-core CPU leg run before this process touches the GPU
    traffic = None
    if not args.no_pmc and not args.adjoint:
        traffic = measure_traffic(args)
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

    fl, fl_src = flops.workload_flops(args.workload), "counted"
    if fl is None:
        # cloudy (g, layer) cells of this workload, for the routine-by-routine tally
        mc = api.mcica_subcol_sw(1, cols.ncol, cols.nlay, cols.icld, cols.permuteseed, cols.irng,
                                 d["play"], d["cldfrac"], d["ciwp"], d["clwp"], d["rei"],
                                 d["rel"], d["tauc"], d["ssac"], d["asmc"], d["fsfc"], d["alpha"])
        cloudy = float(mc["cldfmcl"].sum().item())
        del mc
        torch.cuda.synchronize()
        fl, fl_src = flops.call_flops(cols.ncol, cols.nlay, cloudy), "tally"

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
    host64 = {k: api.to_host(v) for k, v in o64.items()}
    hostmx = {k: api.to_host(v) for k, v in omx.items()}
    achieved = fl / (kern_ms * 1e-3) / 1e12
    kname = f"rt_kernel<{0 if args.precision == 'fp64' else 1}, 1>"
    if traffic is not None:
        rw = {"read": traffic["read"], "write": traffic["write"]}
        t_src = (f"measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over the timed "
                 f"call, scaled x{traffic['read_scale']:.3f} / x{traffic['write_scale']:.3f} by "
                 f"the same counters on a 1 GiB elementwise kernel")
    elif args.workload == "mcica8192x72":
        rw = RECORDED_TRAFFIC[args.precision]
        t_src = "recorded: profiles/pmc_rt_kernel_mcica8192x72.md (this run could not measure it)"
    else:
        rw, t_src = None, "not measured"
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
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS,
                     "traffic": None if rw is None else rw["read"] + rw["write"],
                     "traffic_read_write": rw, "traffic_source": t_src,
                     "traffic_over_hbm_peak_at_event_time":
                         None if rw is None else (rw["read"] + rw["write"]) / (kern_ms * 1e-3)
                         / (PEAK_HBM_GBS * 1e9),
                     "kernel": kname,
                     "note": "FP64 vector-ALU work, no matrix-core (MFMA) contraction: algorithmic "
                             "FLOPs over the per-step HIP-event time on the launch stream, "
                             "against MI355X's FP64 vector peak (78.6 TFLOP/s, equal to its FP64 "
                             "matrix peak)",
                     "flops_per_step": fl, "flops_source": fl_src + (
                         ": +, -, *, / and sqrt executed by the oracle's operation-counting build "
                         "on the whole workload (rrtmg_sw_amd/data/flops_counted.json)"
                         if fl_src == "counted" else ": rrtmg_sw_amd/flops.py"),
                     "event_ms_per_step": kern_ms},
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
    del out, out_o, o64, omx
    line["adjoint"] = adjoint_ratio(args)
    line["nomcica"] = nomcica_form(args)
    if args.workload == "mcica8192x72":
        line["mt"] = mt_form(cols, d)
        del d
        torch.cuda.empty_cache()
        line["host_abi"] = host_abi_form(cols)
        torch.cuda.empty_cache()
        line["adjoint_large"] = adjoint_large(args)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        line["cpu_baseline"], ref = cpu_baseline(cols)
        n = ref["swuflx"].shape[0]
        line["forms"]["max_abs_error_vs_cpu_reference"] = {
            "columns": n, "reference": line["cpu_baseline"]["kind"],
            "fp64": {k: float(abs(host64[k][:n] - ref[k]).max()) for k in ref},
            "mixed": {k: float(abs(hostmx[k][:n] - ref[k]).max()) for k in ref}}
    if cpu_all is not None:
        line["cpu_baseline_all_cores"] = cpu_all
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
