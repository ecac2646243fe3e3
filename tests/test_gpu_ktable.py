"""Tables read from a k-table file (rrtmgsw_init_nc, the src/rrtmg_sw_read_nc.f90 build) drive
the GPU path exactly as the compiled-in tables do: a file holding the same data gives bit-identical
fluxes and heating rates, a file with changed data gives changed ones."""
import os
import sys

import numpy as np
import pytest

from rrtmg_sw_amd.workloads import make_columns
from test_gpu_parity import mc_args, sw_args

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import make_ktable_nc as mk  # noqa: E402

pytestmark = pytest.mark.gpu

OUT = ("swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc")


def _run(api, cols):
    mc = api.mcica_subcol_sw(**mc_args(cols))
    return api.rrtmg_sw(**sw_args(cols, mc))


def test_ktable_file_drives_the_gpu_path(gpu_lib, tmp_path, monkeypatch):
    api = gpu_lib
    cols = make_columns(16, 72, icld=2, iaer=10, seed=910)
    ref = _run(api, cols)
    same = str(tmp_path / "rrtmg_sw.nc")
    mk.write(same, version=2)
    arrays = mk.build_arrays()
    # bands 18 and 19 (water vapour, 4 000-5 150 cm-1) absorb twice as much below the tropopause
    arrays["KeySpeciesAbsorptionCoefficientsLowerAtmos"][:, :, :, :, 2:4, 0] *= 2.0
    changed = str(tmp_path / "changed.nc")
    from scipy.io import netcdf_file
    f = netcdf_file(changed, "w")
    for d, n in mk.DIMS.items():
        f.createDimension(d, n)
    for v, a in arrays.items():
        f.createVariable(v, "d", tuple(reversed(mk.VARS[v])))[...] = a.T
    f.close()
    try:
        api.rrtmg_sw_finalize()
        monkeypatch.chdir(tmp_path)
        api.rrtmg_sw_ini(1004.64, ktable_file="")       # rrtmg_sw.nc in the working directory
        with pytest.raises(api.RRTMGSWError) as e:      # one table source per process
            api.rrtmg_sw_ini(1004.64)
        assert "rrtmgsw_finalize" in e.value.message
        api.rrtmg_sw_ini(1004.64, ktable_file="rrtmg_sw.nc")
        got = _run(api, cols)
        for k in OUT:
            assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k
        api.rrtmg_sw_finalize()
        api.rrtmg_sw_ini(1004.64, ktable_file=changed)
        alt = _run(api, cols)
        d = np.abs(np.asarray(alt["swdflx"]) - np.asarray(ref["swdflx"]))
        print("doubled band-18/19 lower-atmosphere absorption: max swdflx change %.3f W/m2" % d.max())
        assert d.max() > 0.5
        # a file that fails to read leaves the library uninitialised, free to start again
        api.rrtmg_sw_finalize()
        with pytest.raises(api.RRTMGSWError) as e:
            api.rrtmg_sw_ini(1004.64, ktable_file=str(tmp_path / "missing.nc"))
        assert "cannot open" in e.value.message
        with pytest.raises(api.RRTMGSWError) as e:
            _run(api, cols)
        assert e.value.code == 1        # RRTMGSW_ERR_NOT_INITIALIZED
    finally:
        api.rrtmg_sw_finalize()
        api.rrtmg_sw_ini(1004.64)
    assert all(np.array_equal(np.asarray(_run(api, cols)[k]), np.asarray(ref[k])) for k in OUT)
