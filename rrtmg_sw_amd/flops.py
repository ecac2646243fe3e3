"""Algorithmic floating-point operation count of the GCM-mode shortwave path.

workload_flops(name) is the counted figure: the +, -, *, / and sqrt that the oracle executes on the
whole workload (mcica_subcol_sw + rrtmg_sw), counted by its operation-counting build and stored
in data/flops_counted.json by oracle/count_flops.py. The tally below is an independent reading of
the reference's sources, routine by routine, kept as a cross-check (they agree within 10 %) and
for workloads the table does not hold.

The tally is the reference algorithm's own work (not what a particular implementation issues):
each +, -, *, / and sqrt counts 1; comparisons, min/max, table gathers, integer and random-number
arithmetic count 0. It is tallied routine by routine from the reference sources:

per (column, g-point, layer), every layer (281 in total):
  taumol      25   key-species 8-corner interpolation (15), continua (10), species ratio share
                   (e.g. src/rrtmg_sw_taumol.f90:286-330; averaged over the 14 band forms)
  optics      19   clear-sky optics and delta scaling (src/rrtmg_sw_spcvmc.f90:442-496)
  total-sky   11   clear + cloud combination, computed for every layer (:500-505)
  reftra     104   general two-stream branch (src/rrtmg_sw_reftra.f90:164-322, two table
                   exponentials at 4 each)
  direct      28   direct beam and clear/cloud blending (:539-587)
  vrtqdr      86   up sweep 14, down sweep 15, level fluxes 14, for each of the two skies
                   (src/rrtmg_sw_vrtqdr.f90:122-169)
  accumulate   8   four live g-point sums (src/rrtmg_sw_spcvmc.f90:607-622)
per cloudy sub-column cell (cldfmc > 0), in addition (164):
  cldprmc     60   Fu ice + Hu-Stamnes liquid optics and delta-M (src/rrtmg_sw_cldprmc.f90:235-337)
  reftra     104   total-sky two-stream solution (src/rrtmg_sw_spcvmc.f90:534-537)
per column and layer: setcoef + inatm 50, heating rates 6; per column and g-point: 7.
"""
PER_GLAYER = 25 + 19 + 11 + 104 + 28 + 86 + 8
PER_CLOUDY_CELL = 60 + 104
PER_LAYER = 50 + 6
PER_G = 7
NGPT = 112


def column_flops(nlay, cloudy_cells=0.0):
    """Algorithmic FLOPs of one column of `nlay` layers with `cloudy_cells` cloudy
    (g-point, layer) cells."""
    return NGPT * nlay * PER_GLAYER + cloudy_cells * PER_CLOUDY_CELL + nlay * PER_LAYER + NGPT * PER_G


def call_flops(ncol, nlay, cloudy_cells_total):
    return ncol * column_flops(nlay) + cloudy_cells_total * PER_CLOUDY_CELL


def fused_bytes_per_column(nlay, iaer=0):
    """Compulsory HBM bytes of one column for the fused device-resident call: inputs read once
    (profiles, the mcica_subcol_sw physical cloud inputs, aerosol), outputs written once."""
    L = nlay
    prof = 8 * L + 2 * (L + 1) + 6           # play tlay h2o o3 co2 ch4 n2o o2, plev tlev, sfc/albedo/mu
    cloud = 7 * L                            # cldfrac ciwp clwp rei rel alpha + tauc-ssac-asmc-fsfc of 0
    aer = 42 * L if iaer == 10 else 0
    out = 4 * (L + 1) + 2 * L
    return 8 * (prof + cloud + aer + out)


_COUNTED = None


def workload_flops(name):
    """Counted FLOPs of one call on the named workload, or None when it was not counted."""
    global _COUNTED
    if _COUNTED is None:
        import json
        import os
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "flops_counted.json")
        with open(path) as f:
            _COUNTED = json.load(f)
    rec = _COUNTED.get(name)
    return None if rec is None else rec["total"]
