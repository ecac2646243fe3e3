"""rrtmg_sw_amd -- RRTMG_SW's GCM-mode McICA shortwave path on AMD Instinct MI355X (gfx950).

The compute lives in librrtmgsw_hip.so (hand-written HIP kernels behind a C-ABI, see
include/rrtmgsw.h); this package holds the Python mirror of the reference interface
(rrtmg_sw_amd.api), synthetic workloads (rrtmg_sw_amd.workloads) and the build recipe
(rrtmg_sw_amd.build). Importing the package does not touch the GPU.
"""
from . import workloads  # noqa: F401

__all__ = ["api", "workloads", "build"]
