"""The column-mode program's readers (rrtmg_sw_amd/column/rrtmgsw_column.cpp, run with --dump,
no GPU work) against the values tests/golden/make_column_kats.py read from the same input
files of the reference's example suite (INPUT_RRTM records for IATM = 0, IN_CLD_RRTM,
IN_AER_RRTM), and its refusals of what it does not cover."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
EXE = os.path.join(ROOT, "rrtmg_sw_amd", "column", "rrtmgsw_column")
INPUTS = os.path.join(HERE, "golden", "column_inputs")
KATS = np.load(os.path.join(HERE, "golden", "column_kats.npz"))

pytestmark = pytest.mark.skipif(not os.path.exists(EXE), reason="column program not built")


def dump(case_dir):
    r = subprocess.run([EXE, "--dump", case_dir], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, *v = line.split()
        out[k] = np.array([float(x) for x in v])
    return out


@pytest.mark.parametrize("name", [str(n) for n in KATS["cases"]])
def test_readers_match(name):
    g = lambda k: KATS[f"{name}/{k}"]  # noqa: E731
    d = dump(os.path.join(INPUTS, name))
    L = int(g("nlay"))
    iaer, iout, icld, idelm, juldat, isolvar, scon, solcycfrac, zenith = d["scalars"]
    assert (iaer, iout, idelm, juldat, isolvar) == (g("iaer"), g("iout"), g("idelm"), g("juldat"), g("isolvar"))
    assert (scon, solcycfrac) == (g("scon"), g("solcycfrac"))
    assert zenith == g("zenith")
    for k in ("pavel", "tavel", "pz", "coldry", "semiss", "bndsolvar", "indsolvar"):
        assert np.array_equal(d[k], np.ravel(g(k))), k
    assert np.array_equal(d["wkl"], np.ravel(g("wkl"), order="F"))
    assert np.array_equal(d["overlap"], [g("imca"), g("idcor"), g("decorr_con"), g("lat"), g("juldat_alpha")])
    assert np.array_equal(d["altz"], g("altz"))
    if f"{name}/inflag" in KATS:
        assert icld >= 1
        assert np.array_equal(d["cldflags"], [g("inflag"), g("iceflag"), g("liqflag")])
        for k in ("cldfrac", "ciwp", "clwp", "rei", "rel"):
            assert np.array_equal(d[k], g(k)), k
    if iaer == 10:
        for k in ("tauaer", "ssaaer", "asmaer"):
            np.testing.assert_allclose(d[k], np.ravel(g(k), order="F"), rtol=1e-15, atol=0, err_msg=k)
        assert L == 51 and d["tauaer"].max() > 0


def test_refuses_generated_atmosphere(tmp_path):
    """IATM = 1 (layers built by src/rrtatm.f) is not covered: a clear error, no output."""
    src = os.path.join(INPUTS, "MLS-clr")
    for f in os.listdir(src):
        shutil.copyfile(os.path.join(src, f), tmp_path / f)
    lines = open(tmp_path / "INPUT_RRTM").read().split("\n")
    i = next(k for k, l in enumerate(lines) if l.startswith("$")) + 1
    rec = lines[i].ljust(100)
    assert rec[49] == "0"                 # record 1.2, column 50: iatm
    lines[i] = rec[:49] + "1" + rec[50:]
    open(tmp_path / "INPUT_RRTM", "w").write("\n".join(lines))
    r = subprocess.run([EXE, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "IATM = 1" in r.stderr
    assert not (tmp_path / "OUTPUT_RRTM").exists()
