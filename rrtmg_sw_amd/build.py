"""Build librrtmgsw_hip.so (HIP kernels for gfx950 + C-ABI host code + embedded tables) in-tree.

    python -m rrtmg_sw_amd.build        # or rrtmg_sw_amd.build.build()

Objects go to rrtmg_sw_amd/_build/; the library to rrtmg_sw_amd/librrtmgsw_hip.so; the
column-mode program to rrtmg_sw_amd/column/rrtmgsw_column. The Fortran iso_c_binding modules
(rrtmg_sw_amd/fortran/) are built as well when amdflang is present.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "librrtmgsw_hip.so")
BLOB = os.path.join(HERE, "data", "rrtmg_sw_tables.bin")
ARCH = os.environ.get("RRTMGSW_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
# Packed FP32 instructions are disabled: the compiler emits them from ordinary float code, and
# some of their forms have given wrong results on this hardware without any error.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I", os.path.join(ROOT, "include")]
SOURCES = ["rrtmgsw_kernels.hip", "rrtmgsw_adjoint.hip", "rrtmgsw_api.cpp", "rrtmgsw_tables.cpp",
           "rrtmgsw_mtjump.cpp", "rrtmgsw_nc.cpp"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd[:3])} ...")
    return r


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    for src in SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") or src == "rrtmgsw_api.cpp" else []
        _run([HIPCC, *HIP_FLAGS, *lang, "-c", os.path.join(CSRC, src), "-o", obj], verbose)
        objs.append(obj)
    blob_s = os.path.join(BUILD, "rrtmgsw_blob.S")
    with open(blob_s, "w") as f:
        f.write(".section .rodata\n.balign 64\n.global rrtmgsw_blob_start\n.global rrtmgsw_blob_end\n"
                f"rrtmgsw_blob_start:\n.incbin \"{BLOB}\"\nrrtmgsw_blob_end:\n"
                ".section .note.GNU-stack,\"\",@progbits\n")
    blob_o = os.path.join(BUILD, "rrtmgsw_blob.o")
    _run(["gcc", "-c", blob_s, "-o", blob_o], verbose)
    objs.append(blob_o)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs], verbose)
    os.replace(tmp, LIB)
    # the column-mode program (C++ host code over the C-ABI)
    col = os.path.join(HERE, "column")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"), "-o",
          os.path.join(col, "rrtmgsw_column"), os.path.join(col, "rrtmgsw_column.cpp"), f"-L{HERE}",
          "-lrrtmgsw_hip", "-Wl,-rpath,$ORIGIN/.."], verbose)
    fortran = os.path.join(HERE, "fortran", "build_fortran.sh")
    if os.path.exists(fortran) and shutil.which("amdflang"):
        _run(["bash", fortran], verbose)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
