"""Host-side MT19937 jump-ahead (rrtmg_sw_amd/csrc/rrtmgsw_mtjump.cpp: characteristic polynomial by
Berlekamp-Massey, x^J mod P by squaring) against sequential generation, on the CPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(os.path.dirname(HERE), "rrtmg_sw_amd", "csrc", "rrtmgsw_mtjump.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_jump_ahead_matches_sequential_stream(tmp_path):
    exe = str(tmp_path / "mtjump_check")
    subprocess.run(["g++", "-O2", "-o", exe, os.path.join(HERE, "mtjump_check.cpp"), SRC], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("mismatches=0") == 16   # 9 single jumps, 7 multiples
