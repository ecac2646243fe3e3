"""The Fortran drop-in: a host program compiled against the modules rrtmg_sw_init,
mcica_subcol_gen_sw and rrtmg_sw_rad (same routine names and dummies as the reference) calls
mcica_subcol_sw + rrtmg_sw on the GPU; a second one, compiled against the non-McICA driver's
modules, calls its rrtmg_sw. Outputs are checked against the oracle, also with every array held in
one of pcols = ncol + pad columns whose extra columns hold junk (a GCM's pcols-sized arrays): the
drop-in reads the leading ncol columns, as the reference does."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_bind as ob
from rrtmg_sw_amd.workloads import make_columns, overcast

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "rrtmg_sw_amd", "fortran", "example_gcm")
EXE_NM = os.path.join(ROOT, "rrtmg_sw_amd", "fortran", "example_gcm_nomcica")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran example not built")]

ORDER = ["play", "plev", "tlay", "tlev", "tsfc", "h2ovmr", "o3vmr", "co2vmr", "ch4vmr", "n2ovmr",
         "o2vmr", "asdir", "asdif", "aldir", "aldif", "coszen", "cldfrac", "ciwp", "clwp", "rei",
         "rel", "tauc", "ssac", "asmc", "fsfc", "alpha", "tauaer", "ssaaer", "asmaer", "ecaer"]


def _write_inputs(path, cols, icld, iaer):
    hdr = np.array([cols.ncol, cols.nlay, icld, iaer, cols.permuteseed, cols.irng, cols.dyofyr,
                    cols.isolvar, cols.inflgsw, cols.iceflgsw, cols.liqflgsw], dtype=np.int32)
    with open(path, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([cols.adjes, cols.scon], dtype=np.float64).tobytes())
        for k in ORDER:
            f.write(np.asarray(getattr(cols, k), dtype=np.float64).ravel(order="F").tobytes())


def _read_outputs(path, cols):
    raw = open(path, "rb").read()
    ic, ia = np.frombuffer(raw[:8], dtype=np.int32)
    vals = np.frombuffer(raw[8:], dtype=np.float64)
    N, L = cols.ncol, cols.nlay
    sizes = [N * (L + 1), N * (L + 1), N * L, N * (L + 1), N * (L + 1), N * L]
    got, o = {}, 0
    for k, n in zip(NAMES, sizes):
        got[k] = vals[o:o + n].reshape((N, -1), order="F")
        o += n
    return int(ic), int(ia), got


NAMES = ["swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"]


@pytest.mark.parametrize("icld,iaer,pad", [(2, 6, 0), (0, 0, 0), (2, 6, 7)])
def test_fortran_host_program_nomcica(tmp_path, icld, iaer, pad):
    """The non-McICA driver's modules (rrtmg_sw_hip_nomcica.f90): same names as the reference's
    src/rrtmg_sw_rad.nomcica.f90; layer clouds, overcast only."""
    if not os.path.exists(EXE_NM):
        pytest.skip("Fortran example not built")
    cols = overcast(make_columns(20, 45, icld=2, iaer=0, seed=80 + icld))
    cols.ecaer = np.asfortranarray(np.random.default_rng(8).uniform(0.0, 0.2, cols.ecaer.shape))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_inputs(fin, cols, icld, iaer)
    r = subprocess.run([EXE_NM, str(fin), str(fout), str(pad)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    ic, ia, got = _read_outputs(fout, cols)
    ref, ic_r, ia_r = ob.run_nomcica(cols, icld=icld, iaer=iaer)
    assert (ic, ia) == (ic_r, ia_r)
    for k in NAMES:
        d = np.abs(got[k] - ref[k]).max()
        assert d <= (0.01 if k.startswith("swhr") else 0.1), (k, d)


def test_fortran_nomcica_partial_cloud_stops(tmp_path):
    """A partial cloud ends the program with the reference's STOP message."""
    if not os.path.exists(EXE_NM):
        pytest.skip("Fortran example not built")
    cols = overcast(make_columns(4, 30, icld=2, iaer=0, seed=81))
    f = cols.cldfrac.copy()
    f[2, np.argmax(f[2] > 0)] = 0.3
    cols.cldfrac = np.asfortranarray(f)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    _write_inputs(fin, cols, 2, 0)
    r = subprocess.run([EXE_NM, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "PARTIAL CLOUD NOT ALLOWED" in r.stdout + r.stderr


@pytest.mark.parametrize("icld,iaer,pad", [(2, 0, 0), (5, 10, 0), (0, 0, 0), (2, 6, 5),
                                           (5, 10, 3)])
def test_fortran_host_program(tmp_path, icld, iaer, pad):
    cols = make_columns(20, 45, icld=icld, iaer=iaer, seed=70 + icld)
    if iaer == 6:
        cols.ecaer = np.asfortranarray(np.random.default_rng(9).uniform(0.0, 0.2, cols.ecaer.shape))
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    hdr = np.array([cols.ncol, cols.nlay, icld, iaer, cols.permuteseed, cols.irng, cols.dyofyr,
                    cols.isolvar, cols.inflgsw, cols.iceflgsw, cols.liqflgsw], dtype=np.int32)
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([cols.adjes, cols.scon], dtype=np.float64).tobytes())
        for k in ORDER:
            f.write(np.asarray(getattr(cols, k), dtype=np.float64).ravel(order="F").tobytes())
    r = subprocess.run([EXE, str(fin), str(fout), str(pad)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    ic, ia = np.frombuffer(raw[:8], dtype=np.int32)
    vals = np.frombuffer(raw[8:], dtype=np.float64)
    N, L = cols.ncol, cols.nlay
    sizes = [N * (L + 1), N * (L + 1), N * L, N * (L + 1), N * (L + 1), N * L]
    names = ["swuflx", "swdflx", "swhr", "swuflxc", "swdflxc", "swhrc"]
    got, o = {}, 0
    for k, n in zip(names, sizes):
        got[k] = vals[o:o + n].reshape((N, -1), order="F")
        o += n
    mc, _ = ob.run_mcica(cols) if icld else (ob.empty_mcica(cols), 0)
    ref = ob.run_rrtmg(cols, mc, icld=min(icld, 3))
    assert ic == min(icld, 3) if icld <= 3 else ic == 2
    for k in names:
        d = np.abs(got[k] - ref[k]).max()
        assert d <= (0.01 if k.startswith("swhr") else 0.1), (k, d)


def test_fortran_host_program_ktable_file(tmp_path):
    """rrtmg_sw_ini(cpdair, ktable_file=...) of the Fortran module: a k-table file holding the
    compiled-in data gives the same bits as the compiled-in tables."""
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_ktable_nc as mk
    cols = make_columns(12, 40, icld=2, iaer=10, seed=77)
    fin = tmp_path / "in.bin"
    hdr = np.array([cols.ncol, cols.nlay, 2, 10, cols.permuteseed, cols.irng, cols.dyofyr,
                    cols.isolvar, cols.inflgsw, cols.iceflgsw, cols.liqflgsw], dtype=np.int32)
    with open(fin, "wb") as f:
        f.write(hdr.tobytes())
        f.write(np.array([cols.adjes, cols.scon], dtype=np.float64).tobytes())
        for k in ORDER:
            f.write(np.asarray(getattr(cols, k), dtype=np.float64).ravel(order="F").tobytes())
    nc = tmp_path / "rrtmg_sw.nc"
    mk.write(str(nc))
    outs = []
    for extra in ([], [str(nc)]):
        fout = tmp_path / f"out{len(extra)}.bin"
        r = subprocess.run([EXE, str(fin), str(fout), "0", *extra], capture_output=True, text=True,
                           timeout=120)
        assert r.returncode == 0, r.stdout + r.stderr
        outs.append(open(fout, "rb").read())
    assert len(outs[0]) > 8 + 8 * 12 * 40 and outs[0] == outs[1]
