#!/usr/bin/env python3
"""Throughput of the GCM-mode McICA shortwave path on one MI355X.

Workload (one step): mcica_subcol_sw + rrtmg_sw on 8192 columns x 72 layers, McICA maximum-random
overlap (icld = 2) with the KISS generator, synthetic inputs of rrtmg_sw_amd.workloads
"mcica8192x72", all inputs resident in HBM before the timed region. The timed call is the fused
device entry point (rrtmgsw_mcica_sw_dev, FP64). Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--precision fp64|mixed] [--no-cpu]

Both precision forms are timed in the same run; `value` is the one --precision names (FP64 by
default), and `forms` carries both, their speed ratio and the mixed form's largest error
against the FP64 form (fluxes W/m2, heating rates K/day).
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6   # MI355X FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


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


def main():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--precision", default="fp64", choices=["fp64", "mixed"])
    p.add_argument("--workload", default="mcica8192x72")
    p.add_argument("--no-cpu", action="store_true")
    args = p.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has one GPU; the path shards "
              f"by columns with no collective (see DESIGN.md). --gpus {args.gpus} is not run.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from rrtmg_sw_amd import api, flops
    from rrtmg_sw_amd.workloads import workload

    api.rrtmg_sw_ini(1004.64)
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
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel": f"rt_kernel<{'false' if args.precision == 'fp64' else 'true'}>",
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
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(cols)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
