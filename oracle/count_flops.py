#!/usr/bin/env python3
"""Count the algorithmic floating-point operations of the timed workloads with the oracle's
operation-counting build (oracle/liborc_count.so, from oracle/orc_count.cpp) and write them to
rrtmg_sw_amd/data/flops_counted.json, the table rrtmg_sw_amd/flops.py reads.

TEST INFRASTRUCTURE: run by hand after a change to the oracle or the workloads (the JSON is
committed); tests/test_flops.py recounts the stored 64-column samples.

    python oracle/count_flops.py [workload ...]

Per workload: the operations (+, -, *, /, sqrt) of mcica_subcol_sw and of rrtmg_sw (the McICA
workloads; the non-McICA driver's rrtmg_sw for overcast8192x72) and the exp/log/pow calls, over
all columns in chunks of 256 (the KISS streams are per column, so chunking changes nothing), and
the same over the workload's first 64 columns.
"""
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import oracle_bind as ob  # noqa: E402
from rrtmg_sw_amd.workloads import WORKLOADS, workload  # noqa: E402

OUT = os.path.join(ROOT, "rrtmg_sw_amd", "data", "flops_counted.json")
DEFAULT = ["mcica8192x72", "expran4096x60", "overcast8192x72", "expran65536x137"]


def count(cols, nomcica):
    """(mcica_subcol_sw ops, rrtmg_sw ops, exp/log/pow calls, cloudy cells) of `cols`."""
    lib = ob.orc_count()
    if nomcica:
        lib.orc_fp_reset()
        ob.run_nomcica(cols, which="count")
        ops, trans = ob.fp_counts()
        return 0, ops, trans, 0
    lib.orc_fp_reset()
    mc, _ = ob.run_mcica(cols, "count")
    m_ops, m_trans = ob.fp_counts()
    lib.orc_fp_reset()
    ob.run_rrtmg(cols, mc, which="count")
    r_ops, r_trans = ob.fp_counts()
    return m_ops, r_ops, m_trans + r_trans, int(mc["cldfmcl"].sum())


def count_workload(name, chunk=256):
    full = workload(name)
    nomcica = WORKLOADS[name].get("overcast", False)
    tot = [0, 0, 0, 0]
    for c0 in range(0, full.ncol, chunk):
        part = count(full.subset(slice(c0, min(c0 + chunk, full.ncol))), nomcica)
        tot = [a + b for a, b in zip(tot, part)]
    first = count(full.subset(slice(0, 64)), nomcica)
    keys = ("mcica_subcol_sw", "rrtmg_sw", "exp_log_pow", "cloudy_cells")
    rec = dict(zip(keys, tot), ncol=full.ncol, nlay=full.nlay,
               driver="non-McICA rrtmg_sw" if nomcica else "mcica_subcol_sw + rrtmg_sw")
    rec["total"] = tot[0] + tot[1]
    rec["first64"] = dict(zip(keys, first))
    return rec


def main():
    names = sys.argv[1:] or DEFAULT
    table = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            table = json.load(f)
    for name in names:
        t0 = time.perf_counter()
        table[name] = count_workload(name)
        print(f"{name}: {table[name]['total'] / 1e9:.4f} GFLOP "
              f"({time.perf_counter() - t0:.0f} s)", flush=True)
    with open(OUT, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()
